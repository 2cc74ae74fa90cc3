/*
 * gpu_engine.hip -- host side of the MI355X engine: the drop-in sv_gpu_new /
 * sv_pool_gpu_* entry points (src3d/pool.h:98-124,167; reference driver
 * src3d/gpu.c), device memory, coefficient classes and kernel launches.
 *
 * Design (replacing the reference's worker thread, command queue and
 * per-launch device synchronisation, src3d/gpu.c:786-802,1792-1955):
 *   - one HIP stream per engine, every launch asynchronous; the host waits
 *     only in copyfrom and in the session API's explicit sync;
 *   - material handling is resolved on the host into per-cell coefficient
 *     classes (uint8 when at most 256 distinct classes exist, else uint16),
 *     computed with the same averaging, dispersive flags and PEC rules as
 *     the CPU path, so the device executes the CPU semantics exactly;
 *   - E-side boundary work is fused into ystep_e so that the CPU order
 *     (boundary before source, src3d/svit.c:228-229) holds whichever order
 *     the caller uses for apply_source_estep / boundary_estep.
 */
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <vector>
#include <utility>
#include <unordered_map>
#include <omp.h>
#include "gsvit_internal.h"
#include "engine.h"
#include "kernels.h"
#include "engine_host.h"

#define EPS0 SV_EPSILON_0
#define MU0 SV_MU_0

struct EKey {
    uint32_t w[7];
    bool operator==(const EKey &o) const { return memcmp(w, o.w, sizeof(w)) == 0; }
};
struct HKey {
    uint32_t w[5];
    bool operator==(const HKey &o) const { return memcmp(w, o.w, sizeof(w)) == 0; }
};
template <int N>
static size_t
hash_words(const uint32_t *w)
{
    uint64_t h = 1469598103934665603ull;
    for (int q = 0; q < N; q++) {
        h ^= w[q];
        h *= 1099511628211ull;
    }
    return (size_t)h;
}
struct EKeyHash { size_t operator()(const EKey &k) const { return hash_words<7>(k.w); } };
struct HKeyHash { size_t operator()(const HKey &k) const { return hash_words<5>(k.w); } };

static uint32_t
fbits(double v)
{
    float f = (float)v;
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
}

/* ------------------------------------------------------------------ */
/* coefficient classes                                                  */
/* ------------------------------------------------------------------ */

static void
h_coeffs(const SvPool *mp, int i, int j, int k, double c[5])
{
    /* src3d/yee.c:210-233 and sv_yee_data_get_db (src3d/yee.c:1621-1645) */
    const SvSet *set = mp->set;
    const SvYeeData *d = mp->d;
    double dt = d->dt, mu, sg;
    if (mp->nmat <= 0 && !(set->plan.matmode == SV_MATMODE_FULL || set->plan.matmode == SV_MATMODE_MAGNETIC)) {
        c[0] = 1;
        c[1] = dt / MU0 / d->dx;
        c[2] = dt / MU0 / d->dy;
        c[3] = dt / MU0 / d->dz;
        c[4] = dt / MU0 / d->dx;
        return;
    }
    sv_pool_mu_sigast(mp, i, j, k, &mu, &sg);
    c[0] = (1 - sg * dt / 2 / mu) / (1 + sg * dt / 2 / mu);
    c[1] = (dt / mu / d->dx) / (1 + sg * dt / 2 / mu);
    c[2] = (dt / mu / d->dy) / (1 + sg * dt / 2 / mu);
    c[3] = (dt / mu / d->dz) / (1 + sg * dt / 2 / mu);
    c[4] = c[1];
}

static int
pec_type(const SvPool *mp, int i, int j, int k)
{
    return mp->mats[mp->d->mat->data[i][j][k]].type == SV_MAT_PEC;
}

/* cell inside the design region's per-cell coefficient box */
static int
in_pc_box(const GsvGrid &g, int i, int j, int k)
{
    return g.pc_nx > 0 && i >= g.pc_i0 && i < g.pc_i0 + g.pc_nx && j >= g.pc_j0 && j < g.pc_j0 + g.pc_ny
           && k >= g.pc_k0 && k < g.pc_k0 + g.pc_nz;
}

static void e_key_raw(const Engine *E, int i, int j, int k, EKey *key);

static void
e_key(const Engine *E, int i, int j, int k, EKey *key)
{
    if (in_pc_box(E->g, i, j, k)) {
        memset(key->w, 0, sizeof(key->w));
        key->w[6] = GSV_E_PERCELL;
        return;
    }
    e_key_raw(E, i, j, k, key);
}

static void
e_key_raw(const Engine *E, int i, int j, int k, EKey *key)
{
    const SvPool *mp = E->mp;
    int xt, yt, zt, mi;
    double ca[3], cb[3];
    uint32_t fl = 0;
    mi = sv_pool_get_tbc(mp, i, j, k, &xt, &yt, &zt);
    sv_pool_get_cabs(mp, i, j, k, ca, ca + 1, ca + 2, cb, cb + 1, cb + 2);
    if (xt) fl |= GSV_E_TBCX;
    if (yt) fl |= GSV_E_TBCY;
    if (zt) fl |= GSV_E_TBCZ;
    if (mp->nmat > 0 && xt * yt * zt == 0 && sv_material_is_dispersive(mp->mats[mi].type))
        fl |= GSV_E_ADE | ((uint32_t)E->ade_slot_of_mat[mi] << GSV_E_SLOT_SHIFT);
    if (mi >= 0 && i > 0 && j > 0 && k > 0) {
        /* src3d/yee.c:1139-1237 */
        int p0 = pec_type(mp, i, j, k);
        int pxm = pec_type(mp, i - 1, j, k), pym = pec_type(mp, i, j - 1, k), pzm = pec_type(mp, i, j, k - 1);
        int pxy = pec_type(mp, i - 1, j - 1, k), pyz = pec_type(mp, i, j - 1, k - 1), pxz = pec_type(mp, i - 1, j, k - 1);
        int pxyz = pec_type(mp, i - 1, j - 1, k - 1);
        int pm = mp->mats[mi].type == SV_MAT_PEC;
        if (pm || pxm || pym || pzm || pxy || pyz || pxz || pxyz) {
            if (pxm != p0) fl |= GSV_E_PECY | GSV_E_PECZ;
            if (pym != p0) fl |= GSV_E_PECX | GSV_E_PECZ;
            if (pzm != p0) fl |= GSV_E_PECX | GSV_E_PECY;
            if (pxy && !p0) fl |= GSV_E_PECZ;
            if (pxz && !p0) fl |= GSV_E_PECY;
            if (pyz && !p0) fl |= GSV_E_PECX;
        }
    }
    key->w[0] = fbits(ca[0]); key->w[1] = fbits(ca[1]); key->w[2] = fbits(ca[2]);
    key->w[3] = fbits(cb[0]); key->w[4] = fbits(cb[1]); key->w[5] = fbits(cb[2]);
    key->w[6] = fl;
}

/* per-cell coefficients of the design region's cells, from the current
 * voxel permittivities; the region must be free of dispersive and PEC cells */
int
gsv_build_pctab(Engine *E)
{
    const GsvGrid &g = E->g;
    std::vector<GsvECoef> &t = E->pctab_host;
    int bad = 0;
    t.resize((size_t)g.pc_nx * g.pc_ny * g.pc_nz);
    #pragma omp parallel for schedule(static) reduction(| : bad)
    for (int k = 0; k < g.pc_nz; k++)
        for (int j = 0; j < g.pc_ny; j++)
            for (int i = 0; i < g.pc_nx; i++) {
                EKey key;
                e_key_raw(E, i + g.pc_i0, j + g.pc_j0, k + g.pc_k0, &key);
                GsvECoef &c = t[(size_t)i + (size_t)g.pc_nx * (j + (size_t)g.pc_ny * k)];
                memcpy(c.ca, key.w, 12);
                memcpy(c.cb, key.w + 3, 12);
                c.flags = key.w[6];
                c.pad = 0;
                if ((c.flags & (GSV_E_ADE | GSV_E_PECX | GSV_E_PECY | GSV_E_PECZ))
                    || (c.flags & (GSV_E_TBCX | GSV_E_TBCY | GSV_E_TBCZ)) != (GSV_E_TBCX | GSV_E_TBCY | GSV_E_TBCZ))
                    bad = 1;
            }
    if (bad) {
        fprintf(stderr, "Error: the design region touches dispersive or PEC cells\n");
        return 1;
    }
    return 0;
}

/* classes for every cell; ids in device layout (x fastest, row pitch px).
 * Classes are numbered in two groups, those with late(key) false first: the
 * E kernel tells a dispersive class from its id alone (id >= the group
 * boundary, returned in *nearly) and starts the class's state loads with the
 * plane's field loads instead of after the table lookup. */
template <typename KEY, typename HASH, typename KEYFN, typename LATEFN>
static int
build_classes(Engine *E, KEYFN keyfn, LATEFN late, std::vector<KEY> &table, std::vector<uint8_t> &ids_out,
              int *idbytes, int *nearly)
{
    const int nx = E->g.nx, ny = E->g.ny, nz = E->g.nz, px = E->g.px;
    const size_t ncell = (size_t)px * ny * nz;
    std::vector<uint32_t> tmp(ncell, 0);
    int nthr = omp_get_max_threads();
    std::vector<std::vector<KEY>> lkeys(nthr);

    #pragma omp parallel
    {
        int t = omp_get_thread_num();
        std::unordered_map<KEY, uint32_t, HASH> lm;
        KEY last;
        uint32_t lastid = 0;
        int have_last = 0;
        #pragma omp for schedule(static)
        for (int k = 0; k < nz; k++)
            for (int j = 0; j < ny; j++)
                for (int i = 0; i < nx; i++) {
                    KEY key;
                    uint32_t id;
                    keyfn(i, j, k, &key);
                    if (have_last && key == last)
                        id = lastid;
                    else {
                        auto it = lm.find(key);
                        if (it == lm.end()) {
                            id = (uint32_t)lkeys[t].size();
                            lkeys[t].push_back(key);
                            lm.emplace(key, id);
                        } else
                            id = it->second;
                        last = key;
                        lastid = id;
                        have_last = 1;
                    }
                    tmp[(size_t)i + (size_t)px * j + (size_t)px * ny * k] = ((uint32_t)t << 24) | id;
                }
    }
    std::unordered_map<KEY, uint32_t, HASH> gm;
    std::vector<std::vector<uint32_t>> remap(nthr);
    table.clear();
    for (int t = 0; t < nthr; t++) {
        for (const KEY &key : lkeys[t]) {
            auto it = gm.find(key);
            uint32_t id;
            if (it == gm.end()) {
                id = (uint32_t)table.size();
                table.push_back(key);
                gm.emplace(key, id);
            } else
                id = it->second;
            remap[t].push_back(id);
        }
    }
    if (table.size() > 65536) {
        fprintf(stderr, "Error: %zu distinct coefficient classes exceed the 65536 the engine supports\n", table.size());
        return 1;
    }
    {
        /* stable renumbering: the early group, then the late one */
        std::vector<uint32_t> newid(table.size());
        uint32_t nn = 0;
        for (int pass = 0; pass < 2; pass++)
            for (size_t q = 0; q < table.size(); q++)
                if ((int)late(table[q]) == pass)
                    newid[q] = nn++;
        std::vector<KEY> sorted(table.size());
        int ne = 0;
        for (size_t q = 0; q < table.size(); q++) {
            sorted[newid[q]] = table[q];
            ne += !late(table[q]);
        }
        table.swap(sorted);
        for (auto &r : remap)
            for (uint32_t &v : r)
                v = newid[v];
        *nearly = ne;
    }
    *idbytes = table.size() <= 256 ? 1 : 2;
    ids_out.assign(ncell * (*idbytes), 0);
    #pragma omp parallel for schedule(static)
    for (long long q = 0; q < (long long)ncell; q++) {
        uint32_t v = tmp[q];
        uint32_t id = ((v >> 24) < (uint32_t)nthr && (v & 0xFFFFFF) < remap[v >> 24].size()) ? remap[v >> 24][v & 0xFFFFFF] : 0;
        if (*idbytes == 1)
            ids_out[q] = (uint8_t)id;
        else
            ((uint16_t *)ids_out.data())[q] = (uint16_t)id;
    }
    return 0;
}

static int build_rc_tables(Engine *E);
static int build_class_tables(Engine *E, int nslot);

static int
build_tables(Engine *E)
{
    SvPool *mp = E->mp;
    int n, nslot = 0;
    for (n = 0; n < 256; n++)
        E->ade_slot_of_mat[n] = 0;
    if (mp->nmat > 256) {
        fprintf(stderr, "Error: at most 256 tabulated materials are supported\n");
        return 1;
    }
    E->g.ade = 0;
    E->g.ade_poles = 0;
    E->g.rc = 0;
    {
        int nade = 0, nrc = 0;
        for (n = 0; n < mp->nmat; n++) {
            const int t = mp->mats[n].type;
            nade += t == SV_MAT_ADE;
            nrc += t == SV_MAT_DRUDE || t == SV_MAT_CP || t == SV_MAT_PLRC;
            if (t == SV_MAT_CP3) {
                fprintf(stderr, "Error: material type 3 (CP3) is not supported by the MI355X engine\n");
                return 1;
            }
        }
        if (nade && nrc) {
            fprintf(stderr, "Error: ADE materials (types 5, 7) and recursive-convolution materials (types 2, 4, 6) "
                            "cannot share one pool on the MI355X engine\n");
            return 1;
        }
        if (nrc)
            return build_rc_tables(E) || build_class_tables(E, nrc);
    }
    for (n = 0; n < mp->nmat; n++) {
        const SvMatProp *m = mp->mats + n;
        if (m->type != SV_MAT_ADE)
            continue;
        if (nslot >= GSV_MAX_ADE) {
            fprintf(stderr, "Error: at most %d ADE materials are supported\n", GSV_MAX_ADE);
            return 1;
        }
        GsvAde &a = E->adem_host[nslot];
        double dt = mp->d->dt;
        /* increment form (engine.h), coefficients combined in double */
        if (fabs(m->ade_a0 + m->ade_a1 - 1) > 1e-12) {
            fprintf(stderr, "Error: Drude coefficients of material %d do not satisfy a0 + a1 = 1\n", n);
            return 1;
        }
        a.c0 = (float)m->ade_c0; a.c2 = (float)m->ade_c2; a.c3 = (float)m->ade_c3;
        a.cj = (float)(m->ade_c1 * m->ade_a0);
        a.dj = (float)(-m->ade_a0);
        a.da = (float)m->ade_a2;
        for (int q = 0; q < 3; q++) {
            const double sq = m->ade_bp0[q] + m->ade_bp1[q] - 1;
            a.pcj[q] = (float)(m->ade_c1 * m->ade_bp0[q]);
            a.pcp[q] = (float)(-m->ade_c1 * sq);
            a.pj[q] = (float)(-m->ade_bp0[q]);
            a.ps[q] = (float)sq;
            a.b2[q] = (float)m->ade_bp2[q]; a.b3[q] = (float)m->ade_bp3[q]; a.b4[q] = (float)m->ade_bp4[q];
        }
        a.cbm[0] = (float)((dt / m->epsilon / mp->d->dx) / (1 + m->sigma * dt / 2 / m->epsilon));
        a.cbm[1] = (float)((dt / m->epsilon / mp->d->dy) / (1 + m->sigma * dt / 2 / m->epsilon));
        a.cbm[2] = (float)((dt / m->epsilon / mp->d->dz) / (1 + m->sigma * dt / 2 / m->epsilon));
        /* a pole with zero amplitude and zero frequency and damping contributes nothing */
        a.poles = 0;
        for (int q = 0; q < 3; q++)
            if (m->cp3_a[q] != 0 || m->ade_bp0[q] != -1 || m->ade_bp1[q] != 2)
                a.poles = q < 2 ? 2 : 3;
        a.drude = a.da != 0.f;
        /* the Drude-only form (engine.h) needs cj da = dj c2, true without poles */
        {
            const double cj = m->ade_c1 * m->ade_a0, da = m->ade_a2, dj = -m->ade_a0, c2 = m->ade_c2;
            a.qa = (float)(2 * cj * da + c2);
            a.qb = (float)(cj * da);
            if (a.poles == 0 && fabs(cj * da - dj * c2) > 1e-12 * fabs(cj * da)) {
                fprintf(stderr, "Error: Drude coefficients of material %d do not satisfy cj da = dj c2\n", n);
                return 1;
            }
        }
        if (a.poles > E->g.ade_poles)
            E->g.ade_poles = a.poles;
        E->ade_slot_of_mat[n] = nslot++;
        E->g.ade = 1;
    }
    return build_class_tables(E, nslot);
}

/* the recursive-convolution material table: the coefficients that the
 * reference's branches of src3d/yee.c:266-455, 620-905 form per cell and step,
 * combined in double (type 6 from the planner's susceptibility integrals,
 * src3d/plan.c:2765-2824) */
static int
build_rc_tables(Engine *E)
{
    SvPool *mp = E->mp;
    const double dt = mp->d->dt, sp[3] = { mp->d->dx, mp->d->dy, mp->d->dz };
    int n, nslot = 0;
    if (mp->set->sb.ade_cpml_consistent) {
        fprintf(stderr, "Error: ADE_CPML_CONSISTENT applies to ADE materials only\n");
        return 1;
    }
    for (n = 0; n < mp->nmat; n++) {
        const SvMatProp *m = mp->mats + n;
        if (m->type != SV_MAT_DRUDE && m->type != SV_MAT_CP && m->type != SV_MAT_PLRC)
            continue;
        if (nslot >= GSV_MAX_ADE) {
            fprintf(stderr, "Error: at most %d dispersive materials are supported\n", GSV_MAX_ADE);
            return 1;
        }
        GsvRc &r = E->rcm_host[nslot];
        memset(&r, 0, sizeof(r));
        const double eps = m->epsilon / SV_EPSILON_0;
        r.kind = m->type;
        if (m->type == SV_MAT_DRUDE) {
            /* omega_p / c and c dt in the chi terms, the material's nu in the decay */
            const double w = m->drude_omega_p / SV_LIGHT_SPEED, x = exp(-SV_LIGHT_SPEED * dt);
            const double ch0 = w * w * SV_LIGHT_SPEED * dt - w * w * (1 - x), dch0 = -w * w * (1 - x) * (1 - x);
            r.ca = (float)((eps + dch0) / (eps + ch0));
            r.cs = (float)(1.0 / (eps + ch0));
            for (int c = 0; c < 3; c++)
                r.cb[c] = (float)(dt / (eps + ch0) / SV_EPSILON_0 / sp[0]);
            r.dd = (float)exp(-m->drude_nu * dt);
            r.d0 = (float)(-w * w * x * (1 - x) * (1 - x));
        } else if (m->type == SV_MAT_CP) {
            const double g = m->drude_nu, w = m->drude_omega_p, eg = exp(-g * dt);
            double x0 = -(w * w / g / g) * (1 - eg) + dt * w * w / g;
            r.poles = 2;
            for (int q = 0; q < 2; q++) {
                /* kernel -2i A W e^(-i phi) e^((-G + iW) t): chi0 = f (1 - e^(beta dt)), f its
                 * prefactor over (G - iW); the accumulator's drive is chi0 (1 - e^(beta dt)) */
                const double a = m->cp3_a[q], ph = m->cp3_phi[q], om = m->cp3_omega[q], ga = m->cp3_gamma[q];
                const double nr = -2 * a * om * sin(ph), ni = -2 * a * om * cos(ph), den = ga * ga + om * om;
                const double fr = (nr * ga - ni * om) / den, fi = (ni * ga + nr * om) / den;
                const double er = 1 - exp(-ga * dt) * cos(om * dt), ei = -exp(-ga * dt) * sin(om * dt);
                const double cr = fr * er - fi * ei, ci = fr * ei + fi * er;
                x0 += cr;
                r.q0r[q] = (float)(cr * er - ci * ei);
                r.q0i[q] = (float)(cr * ei + ci * er);
                r.rr[q] = (float)(exp(-ga * dt) * cos(om * dt));
                r.ri[q] = (float)(exp(-ga * dt) * sin(om * dt));
            }
            r.ca = (float)(eps / (eps + x0));
            r.cs = (float)(1.0 / (eps + x0));
            for (int c = 0; c < 3; c++)
                r.cb[c] = (float)(dt / (sp[0] * SV_EPSILON_0 * (eps + x0)));
            r.dd = (float)eg;
            r.d0 = (float)(-(w * w / g / g) * (1 - eg) * (1 - eg));
        } else {
            const double xi0 = m->plrc_d_xi + m->plrc_p_xi[0] + m->plrc_p_xi[1];
            const double chi0 = m->plrc_d_chi + m->plrc_p_chi[0] + m->plrc_p_chi[1];
            const double den = 2 * SV_EPSILON_0 * (eps - xi0 + chi0) + m->sigma * dt;
            r.poles = 2;
            r.ca = (float)((2 * SV_EPSILON_0 * (eps - xi0) - m->sigma * dt) / den);
            r.cs = (float)(2 * SV_EPSILON_0 / den);
            for (int c = 0; c < 3; c++) {
                r.cb[c] = (float)(2 * dt / den / sp[c]);
                r.cbm[c] = (float)((dt / m->epsilon / sp[c]) / (1 + m->sigma * dt / 2 / m->epsilon));
            }
            r.dd = (float)exp(-m->drude_nu * dt);
            r.d0 = (float)m->plrc_d_dxi;
            r.d1 = (float)(m->plrc_d_dchi - m->plrc_d_dxi);
            for (int q = 0; q < 2; q++) {
                /* rotation e^((-G - iW) dt) */
                r.rr[q] = (float)(exp(-dt * m->cp3_gamma[q]) * cos(dt * m->cp3_omega[q]));
                r.ri[q] = (float)(-exp(-dt * m->cp3_gamma[q]) * sin(dt * m->cp3_omega[q]));
                r.q0r[q] = (float)m->plrc_p_dxir[q];
                r.q0i[q] = (float)m->plrc_p_dxii[q];
                r.q1r[q] = (float)(m->plrc_p_dchir[q] - m->plrc_p_dxir[q]);
                r.q1i[q] = (float)(m->plrc_p_dchii[q] - m->plrc_p_dxii[q]);
            }
        }
        if (r.poles > E->g.ade_poles)
            E->g.ade_poles = r.poles;
        E->ade_slot_of_mat[n] = nslot++;
        E->g.ade = 1;
        E->g.rc = 1;
    }
    return 0;
}

/* E and H class tables (after the dispersive material slots are known) */
static int
build_class_tables(Engine *E, int nslot)
{
    SvPool *mp = E->mp;
    std::vector<EKey> ekeys;
    std::vector<HKey> hkeys;
    auto ek = [E](int i, int j, int k, EKey *key) { e_key(E, i, j, k, key); };
    auto hk = [mp](int i, int j, int k, HKey *key) {
        double c[5];
        h_coeffs(mp, i, j, k, c);
        for (int q = 0; q < 5; q++)
            key->w[q] = fbits(c[q]);
    };
    auto e_late = [](const EKey &key) { return (key.w[6] & GSV_E_ADE) != 0; };
    auto h_late = [](const HKey &) { return false; };
    int ne = 0, nh = 0;
    if (build_classes<EKey, EKeyHash>(E, ek, e_late, ekeys, E->ecls_host, &E->idbytes_e, &ne))
        return 1;
    if (build_classes<HKey, HKeyHash>(E, hk, h_late, hkeys, E->hcls_host, &E->idbytes_h, &nh))
        return 1;
    E->g.ade_id0 = ne;
    E->etab.resize(ekeys.size());
    for (size_t q = 0; q < ekeys.size(); q++) {
        GsvECoef &c = E->etab[q];
        memcpy(c.ca, ekeys[q].w, 12);
        memcpy(c.cb, ekeys[q].w + 3, 12);
        c.flags = ekeys[q].w[6];
        c.pad = 0;
    }
    E->htab.resize(hkeys.size());
    for (size_t q = 0; q < hkeys.size(); q++) {
        GsvHCoef &c = E->htab[q];
        memcpy(&c.da, hkeys[q].w, 4);
        memcpy(c.db, hkeys[q].w + 1, 12);
        memcpy(&c.dbc, hkeys[q].w + 4, 4);
        c.pad[0] = c.pad[1] = c.pad[2] = 0;
    }
    if (mp->set->sc.verbose > 0)
        printf("GPU plan: %zu E classes (%d-byte ids), %zu H classes (%d-byte ids), %d dispersive materials\n",
               E->etab.size(), E->idbytes_e, E->htab.size(), E->idbytes_h, nslot);
    return 0;
}

/* ------------------------------------------------------------------ */
/* sv_gpu_new                                                            */
/* ------------------------------------------------------------------ */

static int
engine_env_int(const char *name, int def)
{
    const char *v = getenv(name);
    return v ? atoi(v) : def;
}

extern "C" void *
sv_gpu_new(SvPool *mp, SvSet *set)
{
    int ngpu = 0, q;
    Engine *E;
    if (set->sc.usegpu == 0)
        return NULL;
    if (hipGetDeviceCount(&ngpu) != hipSuccess || ngpu == 0) {
        fprintf(stderr, "Error: no HIP device found for GPU computation\n");
        return NULL;
    }
    E = new Engine();
    E->mp = mp;
    E->set = set;
    E->dev = 0;
    for (q = 0; q < SV_MAX_GPUS && q < ngpu; q++)
        if (set->sc.ugpu[q]) {
            E->dev = q;
            break;
        }
    if (hipSetDevice(E->dev) != hipSuccess) {
        delete E;
        return NULL;
    }
    memset(&E->g, 0, sizeof(E->g));
    E->g.nx = set->sp.xres;
    E->g.ny = set->sp.yres;
    E->g.nz = set->sp.zres;
    E->g.px = (set->sp.xres + 15) / 16 * 16;
    E->g.plane = (long long)E->g.px * E->g.ny;
    E->g.perx = set->sb.bx0 == SV_BOUNDARY_PERIODIC;
    E->g.pery = set->sb.by0 == SV_BOUNDARY_PERIODIC;
    E->g.ade_cpml = set->sb.ade_cpml_consistent != 0;
    E->g.dx = (float)set->sp.dx;
    E->g.dy = (float)set->sp.dy;
    E->g.dz = (float)set->sp.dz;
    /* default launch shape; gsv_autotune / gsv_set_launch change it */
    E->bx_h = E->bx_e = 64;
    E->by_h = E->by_e = 4;
    E->kchunk_h = E->kchunk_e = 16;
    if (mp->src->tsf) {
        SvSourceTSF *t = mp->src->tsf;
        const int perx = set->sb.bx0 == SV_BOUNDARY_PERIODIC, pery = set->sb.by0 == SV_BOUNDARY_PERIODIC;
        if ((!perx && t->box_in > set->sp.xres - 2) || (!pery && t->box_jn > set->sp.yres - 2)
            || t->box_kn > set->sp.zres - 2) {
            fprintf(stderr, "Error: the TSF box must end at least two cells before the far faces\n");
            delete E;
            return NULL;
        }
        if (!(t->ia_theta >= 0 && t->ia_theta < M_PI)) {
            fprintf(stderr, "Error: TSF theta outside [0, pi) is not supported\n");
            delete E;
            return NULL;
        }
    }
    if (set->sa.design && (E->g.perx || E->g.pery)) {
        fprintf(stderr, "Error: design regions are not supported in pools with periodic boundaries\n");
        delete E;
        return NULL;
    }
    if (set->sa.design && (unsigned long long)E->g.plane * (E->g.nz + 2) * 4ull >= (1ull << 32)) {
        fprintf(stderr, "Error: a design region needs field arrays below 4 GiB (%d x %d x %d is too large)\n",
                E->g.nx, E->g.ny, E->g.nz);
        delete E;
        return NULL;
    }
    if (set->sa.design) {
        /* voxel (i,j,k) feeds the averaged coefficients of cells up to (i+1, j+1, k+1) */
        E->g.pc_i0 = set->sa.design_i0;
        E->g.pc_j0 = set->sa.design_j0;
        E->g.pc_k0 = set->sa.design_k0;
        E->g.pc_nx = set->sa.design_in + 1 - set->sa.design_i0;
        E->g.pc_ny = set->sa.design_jn + 1 - set->sa.design_j0;
        E->g.pc_nz = set->sa.design_kn + 1 - set->sa.design_k0;
        if (gsv_build_pctab(E)) {
            delete E;
            return NULL;
        }
    }
    if (build_tables(E)) {
        delete E;
        return NULL;
    }
    set->plan.gpumode = SV_GPUMODE_FULL;
    mp->gpu = E;
    return E;
}

/* ------------------------------------------------------------------ */
/* sv_pool_gpu_init: device allocation and static data                  */
/* ------------------------------------------------------------------ */

static int
kappa_axis(const SvCpmlFace *f0, int on0, const SvCpmlFace *fn, int onn, int n, int e_side, std::vector<float> &rk)
{
    rk.assign(n, 1.f);
    for (int i = 0; i < n; i++) {
        double k = 1;
        if (on0 && i < f0->depth) {
            double v = e_side ? f0->kappae[i] : f0->kappah[i];
            if (v > k) k = v;
        }
        if (onn && i >= n - fn->depth) {
            int r = i - (n - fn->depth);
            double v = e_side ? fn->kappae[r] : fn->kappah[r];
            if (v > k) k = v;
        }
        rk[i] = (float)(1.0 / k);
    }
    return 0;
}

static int
setup_cpml(Engine *E)
{
    SvSet *set = E->set;
    SvBoundary *b = E->mp->bnd;
    const SvCpmlFace *faces[6] = { &b->x0, &b->xn, &b->y0, &b->yn, &b->z0, &b->zn };
    int on[6] = { set->sb.bx0 == SV_BOUNDARY_CPML, set->sb.bxn == SV_BOUNDARY_CPML, set->sb.by0 == SV_BOUNDARY_CPML,
                  set->sb.byn == SV_BOUNDARY_CPML, set->sb.bz0 == SV_BOUNDARY_CPML, set->sb.bzn == SV_BOUNDARY_CPML };
    double dn[3] = { set->sp.dx, set->sp.dy, set->sp.dz };
    GsvGrid &g = E->g;
    std::vector<float> rk;

    kappa_axis(faces[0], on[0], faces[1], on[1], g.nx, 1, rk);
    if (dupload(E, (float **)&g.rkex, rk.data(), rk.size())) return 1;
    kappa_axis(faces[2], on[2], faces[3], on[3], g.ny, 1, rk);
    if (dupload(E, (float **)&g.rkey, rk.data(), rk.size())) return 1;
    kappa_axis(faces[4], on[4], faces[5], on[5], g.nz, 1, rk);
    if (dupload(E, (float **)&g.rkez, rk.data(), rk.size())) return 1;
    kappa_axis(faces[0], on[0], faces[1], on[1], g.nx, 0, rk);
    if (dupload(E, (float **)&g.rkhx, rk.data(), rk.size())) return 1;
    kappa_axis(faces[2], on[2], faces[3], on[3], g.ny, 0, rk);
    if (dupload(E, (float **)&g.rkhy, rk.data(), rk.size())) return 1;
    kappa_axis(faces[4], on[4], faces[5], on[5], g.nz, 0, rk);
    if (dupload(E, (float **)&g.rkhz, rk.data(), rk.size())) return 1;

    for (int f = 0; f < 6; f++) {
        GsvCpmlFace &F = g.f[f];
        memset(&F, 0, sizeof(F));
        if (!on[f])
            continue;
        const SvCpmlFace *h = faces[f];
        int dep = h->depth, axis = f / 2;
        size_t cnt;
        if (dep < 1 || (axis == 0 && dep > g.nx) || (axis == 1 && dep > g.ny) || (axis == 2 && dep > g.nz)) {
            fprintf(stderr, "Error: CPML depth %d does not fit the pool\n", dep);
            return 1;
        }
        F.on = 1;
        F.depth = dep;
        std::vector<float4> ce(dep);
        std::vector<float2> ch(dep);
        for (int q = 0; q < dep; q++) {
            ce[q] = make_float4((float)h->be[q], (float)(h->ce[q] / dn[axis]), (float)h->ce[q], 0.f);
            ch[q] = make_float2((float)h->bh[q], (float)(h->ch[q] / dn[axis]));
        }
        float4 *dce;
        float2 *dch;
        if (dupload(E, &dce, ce.data(), dep) || dupload(E, &dch, ch.data(), dep))
            return 1;
        F.ce = dce;
        F.ch = dch;
        if (axis == 0) cnt = (size_t)dep * g.ny * g.nz;
        else if (axis == 1) cnt = (size_t)g.px * dep * g.nz;
        else cnt = (size_t)g.px * g.ny * dep;
        if (dmalloc(E, &F.pe, cnt) || dmalloc(E, &F.ph, cnt))
            return 1;
        CK(hipMemsetAsync(F.pe, 0, cnt * 8, E->st));
        CK(hipMemsetAsync(F.ph, 0, cnt * 8, E->st));
    }
    return 0;
}


/* the 1-D line of a plane-wave source (TSF or SF) on the device: where a cell
 * reads the line (dcomp over the whole pool), the polarisation factors, the
 * line fields and the waveform */
static int
setup_plane_line(Engine *E, const SvSourceTSF *t, GsvTsf &T)
{
    SvSet *set = E->set;
    double th = t->ia_theta, ph = t->ia_phi, ps = t->ia_psi;
    T.corr = t->corr;
    T.ax = sin(th) * cos(ph);
    T.ay = sin(th) * sin(ph);
    T.az = cos(th);
    /* src3d/source.c:53-107 with i0 = j0 = k0 = 0 and i1, j1, k1 = pool size */
    int X = set->sp.xres, Y = set->sp.yres, Z = set->sp.zres;
    int xo, yo;
    if (ph >= 0 && ph <= M_PI / 2) { xo = 0; yo = 0; }
    else if (ph > M_PI / 2 && ph <= M_PI) { xo = X; yo = 0; }
    else if (ph > M_PI && ph <= 3 * M_PI / 2) { xo = X; yo = Y; }
    else { xo = 0; yo = Y; }
    T.oi = xo;
    T.oj = yo;
    T.ok = (th <= M_PI / 2) ? 0 : Z;
    if (th > M_PI / 2 && ph > M_PI / 2 && ph <= M_PI)
        T.ax = 0; /* rx stays 0 in this quadrant, src3d/source.c:89-91 */
    T.gex = cos(ps) * sin(ph) - sin(ps) * cos(th) * cos(ph);
    T.gey = -cos(ps) * cos(ph) - sin(ps) * cos(th) * sin(ph);
    T.gez = sin(ps) * sin(th);
    T.ghx = sin(ps) * sin(ph) + cos(ps) * cos(th) * cos(ph);
    T.ghy = -sin(ps) * cos(ph) + cos(ps) * cos(th) * sin(ph);
    T.ghz = -cos(ps) * sin(th);
    T.res = t->jp_res;
    double dt = set->plan.dt;
    T.multe = 1.0 / t->jp_epsilon * dt / (EPS0 * t->jp_dx);
    T.multh = 1.0 / t->jp_mu * dt / (MU0 * t->jp_dx);
    T.coef = (dt * SV_LIGHT_SPEED - t->jp_dx) / (dt * SV_LIGHT_SPEED + t->jp_dx);
    T.ndata = t->ndata;
    if (dmalloc(E, &T.je, T.res) || dmalloc(E, &T.jh, T.res) || dmalloc(E, &T.jb, 8))
        return 1;
    CK(hipMemsetAsync(T.je, 0, T.res * 8, E->st));
    CK(hipMemsetAsync(T.jh, 0, T.res * 8, E->st));
    CK(hipMemsetAsync(T.jb, 0, 64, E->st));
    {
        double *w;
        if (dupload(E, &w, t->e, t->ndata))
            return 1;
        T.wave = w;
    }
    return 0;
}

/* scattered-field source (src3d/plan.c:1985-2040, src3d/yee.c:1139-1237):
 * its line, and the cells whose components the PEC rule sets, from the class
 * ids.  Their values are written after the E kernel, which also adds the
 * CPML corrections; the reference sets them before those corrections, so
 * PEC-rule cells inside CPML slabs are refused. */
static int
setup_sf(Engine *E)
{
    SvPool *mp = E->mp;
    SvSourceTSF *t = mp->src->sf;
    GsvGrid &g = E->g;
    memset(&E->sfl, 0, sizeof(E->sfl));
    E->has_sf = 0;
    E->n_sf = 0;
    if (!t)
        return 0;
    if (setup_plane_line(E, t, E->sfl))
        return 1;
    std::vector<int4> cells;
    const GsvCpmlFace *F = g.f;
    int slab = 0;
    for (int k = 0; k < g.nz; k++)
        for (int j = 0; j < g.ny; j++)
            for (int i = 0; i < g.nx; i++) {
                const size_t c = ((size_t)k * g.ny + j) * g.px + i;
                const unsigned id = E->idbytes_e == 1 ? E->ecls_host[c]
                                                      : ((const uint16_t *)E->ecls_host.data())[c];
                const uint32_t fl = E->etab[id].flags;
                const int mask = ((fl & GSV_E_PECX) ? 1 : 0) | ((fl & GSV_E_PECY) ? 2 : 0) | ((fl & GSV_E_PECZ) ? 4 : 0);
                if (!mask)
                    continue;
                if ((F[0].on && i >= 1 && i < F[0].depth) || (F[1].on && i >= g.nx - F[1].depth + 1)
                    || (F[2].on && j >= 1 && j < F[2].depth) || (F[3].on && j >= g.ny - F[3].depth + 1)
                    || (F[4].on && k >= 1 && k < F[4].depth) || (F[5].on && k >= g.nz - F[5].depth + 1))
                    slab = 1;
                cells.push_back(make_int4(i, j, k, mask));
            }
    if (slab) {
        fprintf(stderr, "Error: a scattered-field source with PEC cells inside CPML slabs is not supported\n");
        return 1;
    }
    E->has_sf = 1;
    E->n_sf = (int)cells.size();
    if (E->n_sf && dupload(E, &E->sf_cells, cells.data(), cells.size()))
        return 1;
    return 0;
}

/* Beam profiles at lateral cell position (x, y), src3d/source.c:231-251:
 * the Gaussian exp(-(x-xc)^2/2rx^2) exp(-(y-yc)^2/2ry^2) and the fiber mode
 * exp(-r^2 dx^2/w0^2) inside the cutoff radius, with the mode width w0 from
 * the V number at the centre wavelength (V taken with the permittivity
 * difference, as the reference does) */
static double
beam_gauss(int x, int y, double xc, double yc, double rx, double ry)
{
    return exp(-((double)(x - xc)) * ((double)(x - xc)) / rx / rx / 2.0) * exp(-((double)(y - yc)) * ((double)(y - yc)) / ry / ry / 2.0);
}

static double
beam_fiber(int x, int y, double xc, double yc, double radius, double cutoff, double eps_core, double eps_clad,
           double dx, double lambda)
{
    const double v = 2 * M_PI / lambda * radius * dx * (eps_core - eps_clad);
    const double w0 = (0.65 + 1.619 / v / sqrt(v) + 2.879 / pow(v, 6)) * radius * dx;
    const double rv = (x - xc) * (x - xc) + (y - yc) * (y - yc);
    return rv < cutoff * cutoff ? exp(-rv * dx * dx / w0 / w0) : 0.0;
}

/* TSF face multiplier: the faces' "if Gaussian, else if fiber" at
 * src3d/source.c:558-742, 1482-1649 */
static double
tsf_profile(const SvSet *set, int x, int y)
{
    const SvSetSourceTSF *t = &set->ss.tsf;
    if (t->gaussian)
        return beam_gauss(x, y, t->gaussian_fxpos, t->gaussian_fypos, t->gaussian_rx, t->gaussian_ry);
    if (t->fiber)
        return beam_fiber(x, y, t->fiber_fxpos, t->fiber_fypos, t->fiber_radius, t->fiber_cutoff, t->fiber_epsilon_core,
                          t->fiber_epsilon_cladding, set->sp.dx, set->ss.lambda_center);
    return 1.0;
}

/* focused TSF (src3d/plan.c:1855-1979, src3d/source.c:773-930, 1892-2060):
 * the waves' vacuum lines in one [nw][res] array, the per-wave distance and
 * polarisation table, and dt/dx/eps per E face cell */
static int
setup_tsff(Engine *E)
{
    SvPool *mp = E->mp;
    SvSet *set = E->set;
    const SvSourceTSFF *f = mp->src->tsff;
    GsvTsff &T = E->tsff;
    memset(&T, 0, sizeof(T));
    if (!f)
        return 0;
    E->has_tsff = 1;
    const SvSourceTSF *w0 = f->w;
    T.i0 = w0->box_i0; T.j0 = w0->box_j0; T.k0 = w0->box_k0;
    T.in = w0->box_in; T.jn = w0->box_jn; T.kn = w0->box_kn;
    T.skip[0] = w0->skip_i0; T.skip[1] = w0->skip_in; T.skip[2] = w0->skip_j0;
    T.skip[3] = w0->skip_jn; T.skip[4] = w0->skip_k0; T.skip[5] = w0->skip_kn;
    T.nw = f->nw;
    T.res = w0->jp_res;
    const double dt = set->plan.dt;
    T.hcor = (float)(dt / set->sp.dx / MU0);
    double *je, *jh, *jb, *wave;
    if (dmalloc(E, &je, (size_t)T.nw * T.res) || dmalloc(E, &jh, (size_t)T.nw * T.res) || dmalloc(E, &jb, (size_t)T.nw * 8)
        || dupload(E, &wave, w0->e, (size_t)w0->ndata))
        return 1;
    CK(hipMemsetAsync(je, 0, (size_t)T.nw * T.res * 8, E->st));
    CK(hipMemsetAsync(jh, 0, (size_t)T.nw * T.res * 8, E->st));
    CK(hipMemsetAsync(jb, 0, (size_t)T.nw * 64, E->st));
    T.je = je;
    T.jh = jh;
    std::vector<GsvTsf> lines(T.nw);
    std::vector<GsvTsffWave> waves(T.nw);
    const double cx = set->sp.xres / 2 + f->xshift, cy = set->sp.yres / 2 + f->yshift, cz = set->sp.zres / 2 + f->zshift;
    for (int w = 0; w < T.nw; w++) {
        const SvSourceTSF *t = f->w + w;
        GsvTsf &L = lines[w];
        memset(&L, 0, sizeof(L));
        L.res = T.res;
        L.je = je + (size_t)w * T.res;
        L.jh = jh + (size_t)w * T.res;
        L.jb = jb + (size_t)w * 8;
        L.wave = wave;
        L.ndata = t->ndata;
        L.multe = 1.0 / t->jp_epsilon * dt / (EPS0 * t->jp_dx);
        L.multh = 1.0 / t->jp_mu * dt / (MU0 * t->jp_dx);
        L.coef = (dt * SV_LIGHT_SPEED - t->jp_dx) / (dt * SV_LIGHT_SPEED + t->jp_dx);
        const double th = t->ia_theta, ph = t->ia_phi, ps = t->ia_psi, a = f->acorr[w];
        GsvTsffWave &W = waves[w];
        W.corr = t->corr;
        W.ax = sin(th) * cos(ph);
        W.ay = sin(th) * sin(ph);
        W.az = cos(th);
        W.cx = cx; W.cy = cy; W.cz = cz;
        W.pe[0] = a * (sin(ps) * sin(ph) + cos(ps) * cos(th) * cos(ph));
        W.pe[1] = a * (-sin(ps) * cos(ph) + cos(ps) * cos(th) * sin(ph));
        W.pe[2] = a * (-cos(ps) * sin(th));
        W.ph[0] = a * (cos(ps) * sin(ph) - sin(ps) * cos(th) * cos(ph));
        W.ph[1] = a * (-cos(ps) * cos(ph) - sin(ps) * cos(th) * sin(ph));
        W.ph[2] = a * (sin(ps) * sin(th));
    }
    GsvTsf *dl;
    GsvTsffWave *dw;
    if (dupload(E, &dl, lines.data(), lines.size()) || dupload(E, &dw, waves.data(), waves.size()))
        return 1;
    T.lines = dl;
    T.waves = dw;
    const double ecor = dt / set->sp.dx;
    const int lo[3] = { T.i0, T.j0, T.k0 }, hi[3] = { T.in, T.jn, T.kn };
    for (int face = 0; face < 6; face++) {
        const int ax = face / 2, A = (ax + 1) % 3, B = (ax + 2) % 3;
        const int na = hi[A] - lo[A] + 1, nb = hi[B] - lo[B] + 1;
        std::vector<float> ce((size_t)na * nb);
        for (int b = 0; b < nb; b++)
            for (int a = 0; a < na; a++) {
                int c[3];
                c[A] = lo[A] + a;
                c[B] = lo[B] + b;
                c[ax] = (face % 2) ? hi[ax] : lo[ax];
                double eps, sig;
                sv_pool_epsilon_sigma(mp, c[0], c[1], c[2], &eps, &sig);
                ce[(size_t)a + (size_t)na * b] = (float)(ecor / eps);
            }
        float *d;
        if (dupload(E, &d, ce.data(), ce.size()))
            return 1;
        T.coefE[face] = d;
    }
    return 0;
}

/* external plane source: window geometry, dt/dx/eps per window cell */
static int
setup_ext(Engine *E)
{
    SvSet *set = E->set;
    GsvExt &X = E->ext;
    memset(&X, 0, sizeof(X));
    if (!set->ss.ext.valid)
        return 0;
    E->has_ext = 1;
    X.k = set->ss.ext.k;
    X.ni = set->ss.ext.iextto - set->ss.ext.iextfrom;
    X.nj = set->ss.ext.jextto - set->ss.ext.jextfrom;
    X.ijstart = set->ss.ext.ijstart;
    X.jkstart = set->ss.ext.jkstart;
    X.hcor = (float)(set->plan.dt / set->sp.dx / MU0);
    std::vector<float> ce((size_t)X.ni * X.nj);
    for (int b = 0; b < X.nj; b++)
        for (int a = 0; a < X.ni; a++) {
            double eps, sig;
            sv_pool_epsilon_sigma(E->mp, X.ijstart + a, X.jkstart + b, X.k, &eps, &sig);
            ce[(size_t)a + (size_t)X.ni * b] = (float)(set->plan.dt / set->sp.dx / eps);
        }
    float *d;
    if (dupload(E, &d, ce.data(), ce.size()) || dmalloc(E, &E->ext_px, ce.size()) || dmalloc(E, &E->ext_py, ce.size()))
        return 1;
    X.ce = d;
    X.px = E->ext_px;
    X.py = E->ext_py;
    E->ext_planes.resize(3 * (size_t)set->ss.ext.extxres * set->ss.ext.extyres);
    E->ext_win.resize(2 * ce.size());
    return 0;
}

/* one step's external planes of a family (0: E planes for the H step, 1: H
 * planes for the E step): read on the host, the window uploaded, one launch;
 * a missing file skips the step as the reference does */
static int
run_ext(Engine *E, int family, int hside)
{
    SvSet *set = E->set;
    const int nx = set->ss.ext.extxres, n = nx * set->ss.ext.extyres;
    GsvExt &X = E->ext;
    if (sv_ext_read_planes(set, family, set->sc.step_act, E->ext_planes.data()))
        return 0;
    const size_t w = (size_t)X.ni * X.nj;
    for (int b = 0; b < X.nj; b++)
        for (int a = 0; a < X.ni; a++) {
            const size_t q = (size_t)(set->ss.ext.jextfrom + b) * nx + set->ss.ext.iextfrom + a;
            E->ext_win[(size_t)a + (size_t)X.ni * b] = E->ext_planes[q];
            E->ext_win[w + (size_t)a + (size_t)X.ni * b] = E->ext_planes[(size_t)n + q];
        }
    /* the staging vector is reused next call: copy synchronously */
    CK(hipMemcpyAsync(E->ext_px, E->ext_win.data(), w * sizeof(float), hipMemcpyHostToDevice, E->st));
    CK(hipMemcpyAsync(E->ext_py, E->ext_win.data() + w, w * sizeof(float), hipMemcpyHostToDevice, E->st));
    CK(gsv_launch_ext(&E->g, &X, hside, E->st));
    CK(hipStreamSynchronize(E->st));
    return 0;
}

static int
setup_tsf(Engine *E)
{
    SvPool *mp = E->mp;
    SvSet *set = E->set;
    SvSourceTSF *t = mp->src->tsf;
    GsvTsf &T = E->tsf;
    memset(&T, 0, sizeof(T));
    if (!t)
        return 0;
    E->has_tsf = 1;
    T.i0 = t->box_i0; T.j0 = t->box_j0; T.k0 = t->box_k0;
    T.in = t->box_in; T.jn = t->box_jn; T.kn = t->box_kn;
    if (setup_plane_line(E, t, T))
        return 1;
    double dt = set->plan.dt;
    /* per-face coefficients: activity (skip flags, skip-near-material map), PEC at z0, local epsilon */
    double ecor = dt / set->sp.dx, hcor = dt / set->sp.dx / MU0;
    int skips[6] = { t->skip_i0, t->skip_in, t->skip_j0, t->skip_jn, t->skip_k0, t->skip_kn };
    for (int f = 0; f < 6; f++) {
        int axis = f / 2, far = f % 2, na, nb;
        tsf_face_extent(T, E->g.nx, E->g.ny, axis, &na, &nb);
        std::vector<float> ce((size_t)na * nb, 0.f), chh((size_t)na * nb, 0.f);
        for (int b = 0; b < nb; b++)
            for (int a = 0; a < na; a++) {
                int i, j, k;
                if (axis == 0) { i = far ? T.in : T.i0; j = T.j0 + a; k = T.k0 + b; }
                else if (axis == 1) { i = T.i0 + a; j = far ? T.jn : T.j0; k = T.k0 + b; }
                else { i = T.i0 + a; j = T.j0 + b; k = far ? T.kn : T.k0; }
                if (skips[f] || !sv_tsf_face_active(mp, set, i, j, k))
                    continue;
                if (f == 4 && sv_pool_is_pec(mp, i, j, k))
                    continue;
                double eps, sig;
                sv_pool_epsilon_sigma(mp, i, j, k, &eps, &sig);
                /* the profile is taken at the cell whose line value the face uses:
                 * E faces at the near side one cell outside the box (i0 - 1, j0 - 1),
                 * H faces on the face plane; z faces at (i, j) */
                const int xe = axis == 0 && !far ? i - 1 : i, ye = axis == 1 && !far ? j - 1 : j;
                ce[(size_t)a + (size_t)na * b] = (float)(ecor * tsf_profile(set, xe, ye) / eps);
                chh[(size_t)a + (size_t)na * b] = (float)(hcor * tsf_profile(set, i, j));
            }
        E->tsf_face_on[f] = 0;
        long active = 0;
        for (size_t q = 0; q < ce.size(); q++)
            if (ce[q] != 0.f || chh[q] != 0.f) {
                E->tsf_face_on[f] = 1;
                active++;
            }
        if (set->sc.verbose > 1)
            printf("TSF face %d: %d x %d cells, %ld active\n", f, na, nb, active);
        float *dce, *dch;
        if (dupload(E, &dce, ce.data(), ce.size()) || dupload(E, &dch, chh.data(), chh.size()))
            return 1;
        T.coefE[f] = dce;
        T.coefH[f] = dch;
    }
    return 0;
}

static long long
cell_index(const GsvGrid &g, int i, int j, int k)
{
    return (long long)i + (long long)g.px * j + g.plane * k;
}

static int
setup_sources_and_outputs(Engine *E)
{
    SvPool *mp = E->mp;
    SvSet *set = E->set;
    GsvGrid &g = E->g;
    int n, s;
    /* point sources: per-step table, 0 = no write (src3d/source.c:405-437) */
    E->has_src = mp->src->npnts > 0;
    memset(&E->src, 0, sizeof(E->src));
    if (E->has_src) {
        int ns = mp->src->npnts, nst = set->sc.nsteps;
        std::vector<long long> pos(ns);
        std::vector<float> vals((size_t)nst * ns * 6, 0.f);
        for (n = 0; n < ns; n++) {
            const SvSourcePoint *sp = mp->src->sp + n;
            pos[n] = cell_index(g, sp->i, sp->j, sp->k);
            for (s = 0; s < nst; s++) {
                int p = -1, q;
                if (s < sp->ndata && sp->layered_zpos[s] == s)
                    p = s;
                else
                    for (q = 0; q < sp->ndata; q++)
                        if (sp->layered_zpos[q] == s) { p = q; break; }
                if (p < 0)
                    continue;
                float *v = vals.data() + ((size_t)s * ns + n) * 6;
                v[0] = (float)sp->ex[p]; v[1] = (float)sp->ey[p]; v[2] = (float)sp->ez[p];
                v[3] = (float)sp->hx[p]; v[4] = (float)sp->hy[p]; v[5] = (float)sp->hz[p];
            }
        }
        long long *dp;
        float *dv;
        if (dupload(E, &dp, pos.data(), pos.size()) || dupload(E, &dv, vals.data(), vals.size()))
            return 1;
        E->src.n = ns;
        E->src.nsteps = nst;
        E->src.pos = dp;
        E->src.vals = dv;
        E->src_vals_host.swap(vals);
    }
    /* point traces */
    E->npts = set->so.npnts;
    if (E->npts) {
        std::vector<long long> pos(E->npts);
        for (n = 0; n < E->npts; n++)
            pos[n] = cell_index(g, set->so.pnts[n].i, set->so.pnts[n].j, set->so.pnts[n].k);
        if (dupload(E, &E->d_ppos, pos.data(), pos.size()))
            return 1;
        size_t cnt = (size_t)(set->sc.nsteps + 1) * E->npts * 6;
        if (dmalloc(E, &E->d_rec, cnt))
            return 1;
        CK(hipMemsetAsync(E->d_rec, 0, cnt * 4, E->st));
    }
    return 0;
}

static int
setup_liao_periodic(Engine *E)
{
    SvSet *set = E->set;
    SvPool *mp = E->mp;
    GsvGrid &g = E->g;
    const SvSetBoundary *b = &set->sb;
    int types[6] = { b->bx0, b->bxn, b->by0, b->byn, b->bz0, b->bzn };
    memset(&E->liao, 0, sizeof(E->liao));
    E->has_liao = 0;
    for (int f = 0; f < 6; f++)
        if (types[f] == SV_BOUNDARY_LIAO) {
            E->liao.on[f] = 1;
            E->has_liao = 1;
        }
    if (E->has_liao) {
        double dt = set->plan.dt, d3[3] = { set->sp.dx, set->sp.dy, set->sp.dz };
        int res[3] = { g.nx, g.ny, g.nz };
        for (int axis = 0; axis < 3; axis++) {
            if (!E->liao.on[2 * axis] && !E->liao.on[2 * axis + 1])
                continue;
            int na, nb;
            if (axis == 0) { na = g.ny; nb = g.nz; }
            else if (axis == 1) { na = g.nx; nb = g.nz; }
            else { na = g.nx; nb = g.ny; }
            size_t cnt = (size_t)24 * na * nb;
            if (dmalloc(E, &E->liao.store[axis], cnt))
                return 1;
            CK(hipMemsetAsync(E->liao.store[axis], 0, cnt * 4, E->st));
            for (int far = 0; far < 2; far++) {
                if (!E->liao.on[2 * axis + far])
                    continue;
                std::vector<float> cf((size_t)na * nb * 3);
                for (int a = 0; a < na; a++)
                    for (int bb = 0; bb < nb; bb++) {
                        int pos[3];
                        if (axis == 0) { pos[1] = a; pos[2] = bb; }
                        else if (axis == 1) { pos[0] = a; pos[2] = bb; }
                        else { pos[0] = a; pos[1] = bb; }
                        /* epsilon probe on the face plane; zn uses zres-2 (src3d/boundary.c:657) */
                        pos[axis] = far ? (axis == 2 ? res[axis] - 2 : res[axis] - 1) : 0;
                        double ind = 1;
                        if (set->plan.matmode == SV_MATMODE_FULL || set->plan.matmode == SV_MATMODE_ELECTRIC)
                            ind = sqrt(mp->d->epsilon->data[pos[0]][pos[1]][pos[2]] / EPS0);
                        for (int q = 0; q < 3; q++)
                            cf[((size_t)a * nb + bb) * 3 + q] = (float)((dt * SV_LIGHT_SPEED / ind - d3[q]) / (dt * SV_LIGHT_SPEED / ind + d3[q]));
                    }
                float *dcf;
                if (dupload(E, &dcf, cf.data(), cf.size()))
                    return 1;
                E->liao.coef[2 * axis + far] = dcf;
            }
        }
        E->rezero[0] = b->bx0 == SV_BOUNDARY_CPML;
        E->rezero[1] = b->by0 == SV_BOUNDARY_CPML;
        E->rezero[2] = b->bz0 == SV_BOUNDARY_CPML;
    }
    const SvSetMediumBoundary *m = &set->smb;
    memset(&E->per, 0, sizeof(E->per));
    E->per.x = m->bx0 == SV_BOUNDARY_PERIODIC || m->bxn == SV_BOUNDARY_PERIODIC;
    E->per.y = m->by0 == SV_BOUNDARY_PERIODIC || m->byn == SV_BOUNDARY_PERIODIC;
    E->per.z = m->bz0 == SV_BOUNDARY_PERIODIC || m->bzn == SV_BOUNDARY_PERIODIC;
    E->has_per = E->per.x || E->per.y || E->per.z;
    E->per.x0 = m->bx0pos; E->per.xn = m->bxnpos;
    E->per.y0 = m->by0pos; E->per.yn = m->bynpos;
    E->per.z0 = m->bz0pos; E->per.zn = m->bznpos;
    if (E->has_per) {
        if ((E->per.x && (m->bx0pos < 1 || m->bxnpos > g.nx - 1)) || (E->per.y && (m->by0pos < 1 || m->bynpos > g.ny - 1))
            || (E->per.z && (m->bz0pos < 1 || m->bznpos > g.nz - 1)) || m->bx0pos < 0 || m->by0pos < 0 || m->bz0pos < 0
            || m->bxnpos > g.nx || m->bynpos > g.ny || m->bznpos > g.nz) {
            fprintf(stderr, "Error: periodic medium boundary planes must lie inside the pool\n");
            return 1;
        }
    }
    return 0;
}

/* ring depth of a running DFT: the flush tabulates T * nw * 4 phases in LDS (<= 48 KiB) */
static int
dft_ring_depth(int nw)
{
    int T = 1536 / nw;
    return T > 32 ? 32 : (T < 1 ? 1 : T);
}

/* layered TSF, src3d/plan.c:1600-1685 and src3d/source.c:231-322: the face
 * tables (lateral delay and coefficient of every face cell, E and H sides),
 * then the 1-D model run once for its whole history (the reference runs it at
 * the first H source step; nothing reads it earlier) */
static double
ltsf_delay(const SvSourceLTSF *ls, const SvSet *set, int i, int j)
{
    /* src3d/source.c:286-322: distance from the box corner the wave enters
     * first along the in-plane direction, as a time in steps */
    const double phi = ls->ia_phi;
    int rx, ry;
    if (phi >= 0 && phi <= M_PI / 2.0) { rx = i - ls->box_i0; ry = j - ls->box_j0; }
    else if (phi > M_PI / 2.0 && phi <= M_PI) { rx = i - ls->box_in; ry = j - ls->box_j0; }
    else if (phi > M_PI && phi <= 3.0 * M_PI / 2.0) { rx = i - ls->box_in; ry = j - ls->box_jn; }
    else { rx = i - ls->box_i0; ry = j - ls->box_jn; }
    double radius = sqrt((double)(rx * rx + ry * ry));
    const double alpha = atan2((double)ry, (double)rx);
    radius *= cos(alpha - phi);
    return radius * set->sp.dx * sin(ls->ia_theta) / (SV_LIGHT_SPEED * set->plan.dt);
}

static double
ltsf_profile(const SvSet *set, int x, int y, int h_far_x)
{
    /* src3d/source.c:231-251; the H face at in drops the Gaussian unless the
     * fiber profile is on as well (its "if / if-else" there, :1734-1741) */
    const SvSetSourceLTSF *t = &set->ss.ltsf;
    double gm = 1.0, fm = 1.0;
    if (t->gaussian)
        gm = beam_gauss(x, y, t->gaussian_fxpos, t->gaussian_fypos, t->gaussian_rx, t->gaussian_ry);
    if (t->fiber)
        fm = beam_fiber(x, y, t->fiber_fxpos, t->fiber_fypos, t->fiber_radius, t->fiber_cutoff, t->fiber_epsilon_core,
                        t->fiber_epsilon_cladding, set->sp.dx, set->ss.lambda_center);
    if (h_far_x)
        return t->fiber ? fm : 1.0;
    return t->gaussian ? gm : (t->fiber ? fm : 1.0);
}

static int
setup_ltsf(Engine *E)
{
    SvSet *set = E->set;
    SvPool *mp = E->mp;
    const SvSourceLTSF *ls = mp->src->ltsf;
    GsvLtsf &L = E->ltsf;
    memset(&L, 0, sizeof(L));
    E->has_ltsf = ls != NULL;
    if (!ls)
        return 0;
    const SvSetSourceLTSF *t = &set->ss.ltsf;
    const int res = set->sp.zres;
    if (8 * (size_t)(res + 2) * sizeof(double) > 160 * 1024) {
        fprintf(stderr, "Error: the layered TSF model needs z resolution <= 2046 (one workgroup's LDS)\n");
        return 1;
    }
    L.i0 = ls->box_i0; L.j0 = ls->box_j0; L.k0 = ls->box_k0;
    L.i1 = ls->box_in; L.j1 = ls->box_jn; L.k1 = ls->box_kn;
    const int sk[6] = { t->box_boundary_skipi0, t->box_boundary_skipin, t->box_boundary_skipj0,
                        t->box_boundary_skipjn, t->box_boundary_skipk0, t->box_boundary_skipkn };
    memcpy(L.skip, sk, sizeof(sk));
    L.res = res;
    L.nrows = ls->nrows;
    L.sph = sin(ls->ia_phi);
    L.cph = cos(ls->ia_phi);
    const double ecor = set->plan.dt / set->sp.dx, hcor = set->plan.dt / set->sp.dx / SV_MU_0;
    for (int hs = 0; hs < 2; hs++)
        for (int axis = 0; axis < 3; axis++) {
            const int na = axis == 0 ? L.j1 - L.j0 + 1 : L.i1 - L.i0 + 1;
            const int nb = axis == 2 ? L.j1 - L.j0 + 1 : L.k1 - L.k0 + 1;
            std::vector<double> q(2 * (size_t)na * nb), cf(q.size());
            for (int f = 0; f < 2; f++)
                for (int a = 0; a < na; a++)
                    for (int b = 0; b < nb; b++) {
                        int i, j, k;
                        if (axis == 0) { i = f ? L.i1 : L.i0; j = L.j0 + a; k = L.k0 + b; }
                        else if (axis == 1) { i = L.i0 + a; j = f ? L.j1 : L.j0; k = L.k0 + b; }
                        else { i = L.i0 + a; j = L.j0 + b; k = f ? L.k1 : L.k0; }
                        /* E faces sample outside the near faces, H faces at the face cells */
                        const int si = (!hs && axis == 0 && !f) ? i - 1 : i, sj = (!hs && axis == 1 && !f) ? j - 1 : j;
                        const size_t at = ((size_t)f * na + a) * nb + b;
                        q[at] = ltsf_delay(ls, set, si, sj);
                        /* only the k0 face consults the skip map (src3d/source.c:1073, 1833) */
                        if (axis == 2 && !f && !sv_tsf_face_active(mp, set, i, j, k)) {
                            cf[at] = 0;
                            continue;
                        }
                        if (hs) {
                            cf[at] = hcor * ltsf_profile(set, si, sj, axis == 0 && f);
                        } else {
                            double eps, sig;
                            sv_pool_epsilon_sigma(mp, i, j, k, &eps, &sig);
                            cf[at] = ecor * ltsf_profile(set, si, sj, 0) / eps;
                        }
                    }
            double *p;
            if (dupload(E, &p, q.data(), q.size())) return 1;
            L.q[hs][axis] = p;
            if (dupload(E, &p, cf.data(), cf.size())) return 1;
            L.coef[hs][axis] = p;
        }
    /* the 1-D model, src3d/zpool.c:38-338 */
    GsvLtsfLine M;
    memset(&M, 0, sizeof(M));
    const double zdt = ls->zdt, zdx = ls->zdx, corr = ls->corr, st = sin(ls->ia_theta);
    const double i0 = sqrt(ls->zeps[0]), i1 = sqrt(ls->zeps[res - 1]);
    M.res = res;
    M.nrows = ls->nrows;
    M.ndata = ls->ndata;
    M.multe = zdt / (SV_EPSILON_0 * zdx) / corr;
    M.multh = zdt / (SV_MU_0 * zdx) / corr;
    M.m0 = (zdt * SV_LIGHT_SPEED / i0 - zdx) / (zdt * SV_LIGHT_SPEED / i0 + zdx);
    M.m1 = (zdt * SV_LIGHT_SPEED / i1 - zdx) / (zdt * SV_LIGHT_SPEED / i1 + zdx);
    M.st = st;
    M.inj_e_s = zdt / zdx / (SV_EPSILON_0 * ls->zeps[10]) * sin(ls->ia_psi);
    M.inj_e_c = zdt / zdx / (SV_EPSILON_0 * ls->zeps[10]) * cos(ls->ia_psi);
    M.inj_h_s = zdt / zdx / SV_MU_0 * sin(ls->ia_psi);
    M.inj_h_c = zdt / zdx / SV_MU_0 * cos(ls->ia_psi);
    M.cez = 2 * sqrt(SV_MU_0 / SV_EPSILON_0) * st;
    M.chz = 2 * sqrt(SV_EPSILON_0 / SV_MU_0) * st;
    double *p;
    if (dupload(E, &p, ls->zeps, res)) return 1;
    M.eps = p;
    if (dupload(E, &p, ls->zmu, res)) return 1;
    M.mu = p;
    if (dupload(E, &p, ls->e, ls->ndata)) return 1;
    M.wave = p;
    if (dmalloc(E, &M.hist, 6 * (size_t)M.nrows * res))
        return 1;
    CK(gsv_launch_ltsf_line(&M, E->st));
    CK(hipStreamSynchronize(E->st));
    L.hist = M.hist;
    return 0;
}

/* OUT_FORCE boxes: eps and mu of every face cell in the kernel's order,
 * per-step results zeroed once (each step is added to once) */
static int
setup_forces(Engine *E)
{
    SvSet *set = E->set;
    SvPool *mp = E->mp;
    E->forces.clear();
    for (int n = 0; n < set->so.nforces; n++) {
        const SvOutputForce *o = set->so.forces + n;
        GsvForce f;
        memset(&f, 0, sizeof(f));
        f.i0 = o->box_i0; f.j0 = o->box_j0; f.k0 = o->box_k0;
        f.i1 = o->box_in; f.j1 = o->box_jn; f.k1 = o->box_kn;
        f.ncells = 2 * ((long long)(f.j1 - f.j0) * (f.k1 - f.k0) + (long long)(f.i1 - f.i0) * (f.k1 - f.k0)
                        + (long long)(f.i1 - f.i0) * (f.j1 - f.j0));
        std::vector<double2> em(f.ncells);
        for (long long q = 0; q < f.ncells; q++) {
            int i, j, k, nx, ny, nz;
            double eps, mu, s;
            force_cell_of(f, q, i, j, k, nx, ny, nz);
            sv_pool_epsilon_sigma(mp, i, j, k, &eps, &s);
            sv_pool_mu_sigast(mp, i, j, k, &mu, &s);
            em[q] = make_double2(eps, mu);
        }
        double2 *pe;
        if (dupload(E, &pe, em.data(), em.size()))
            return 1;
        f.em = pe;
        const size_t cnt = 3 * ((size_t)set->sc.nsteps + 1);
        if (dmalloc(E, &f.out, cnt))
            return 1;
        CK(hipMemsetAsync(f.out, 0, cnt * 8, E->st));
        E->forces.push_back(f);
    }
    return 0;
}

/* periodic far field, src3d/pfarfield.c:36-62; bins zeroed, point lists and
 * istart uploaded once */
static int
setup_pfarfield(Engine *E)
{
    SvSet *set = E->set;
    SvPool *mp = E->mp;
    const SvSetPFarfield *sf = &set->spf;
    GsvPFarfield &F = E->pff;
    memset(&F, 0, sizeof(F));
    E->has_pff = sf->nrs > 0;
    if (!E->has_pff)
        return 0;
    F.i0 = sf->box_i0; F.j0 = sf->box_j0; F.k0 = sf->box_k0;
    F.i1 = sf->box_in; F.j1 = sf->box_jn; F.k1 = sf->box_kn;
    F.ix = F.i1 - F.i0; F.iy = F.j1 - F.j0;
    F.pimin = sf->pimin; F.pimax = sf->pimax; F.pjmin = sf->pjmin; F.pjmax = sf->pjmax;
    const long long nimg = (F.pimax > F.pimin && F.pjmax > F.pjmin)
                           ? (long long)(F.pimax - F.pimin) * (F.pjmax - F.pjmin) : 0;
    F.nwork = nimg * 2 * F.ix * F.iy;
    if (F.nwork >= (1LL << 31) * 256) {
        fprintf(stderr, "Error: PERIODIC_NFFF: too many images x plane cells\n");
        return 1;
    }
    int sk[2][5] = { { sf->box_boundary_skipk0, sf->skipk0_imin, sf->skipk0_jmin, sf->skipk0_imax, sf->skipk0_jmax },
                     { sf->box_boundary_skipkn, sf->skipkn_imin, sf->skipkn_jmin, sf->skipkn_imax, sf->skipkn_jmax } };
    memcpy(F.skip, sk, sizeof(sk));
    F.nrs = sf->nrs;
    F.ndata = mp->pfarfield->ndata;
    std::vector<int> ist(F.nrs);
    for (int n = 0; n < F.nrs; n++)
        ist[n] = mp->pfarfield->prpoints[n].istart;
    int *p;
    if (dupload(E, &p, sf->ri, F.nrs)) return 1;
    F.ri = p;
    if (dupload(E, &p, sf->rj, F.nrs)) return 1;
    F.rj = p;
    if (dupload(E, &p, sf->rk, F.nrs)) return 1;
    F.rk = p;
    if (dupload(E, &p, ist.data(), F.nrs)) return 1;
    F.istart = p;
    size_t cnt = (size_t)F.nrs * 3 * F.ndata;
    if (dmalloc(E, &F.bins, cnt))
        return 1;
    CK(hipMemsetAsync(F.bins, 0, cnt * 8, E->st));
    F.dx = set->sp.dx; F.dy = set->sp.dy;
    F.dt = set->plan.dt;
    F.c = SV_LIGHT_SPEED;
    F.post = sf->postprocess != 0;
    F.ppstart = sf->ppstart;
    return 0;
}

static int
setup_farfield(Engine *E)
{
    SvSet *set = E->set;
    SvPool *mp = E->mp;
    SvSetFarfield *sf = &set->sf;
    GsvGrid &g = E->g;
    memset(&E->ff, 0, sizeof(E->ff));
    memset(&E->dft, 0, sizeof(E->dft));
    E->has_ff = sf->nrs > 0;
    E->has_dft = sf->ndft > 0;
    if (!E->has_ff && !E->has_dft)
        return 0;
    GsvFarfieldTD &F = E->ff;
    F.i0 = sf->box_i0; F.j0 = sf->box_j0; F.k0 = sf->box_k0;
    F.i1 = sf->box_in; F.j1 = sf->box_jn; F.k1 = sf->box_kn;
    F.ix = F.i1 - F.i0; F.iy = F.j1 - F.j0; F.iz = F.k1 - F.k0;
    F.ncells = 2LL * ((long long)F.iy * F.iz + (long long)F.ix * F.iz + (long long)F.ix * F.iy);
    if (E->has_ff) {
        int sk[6][5] = {
            { sf->box_boundary_skipi0, sf->skipi0_jmin, sf->skipi0_kmin, sf->skipi0_jmax, sf->skipi0_kmax },
            { sf->box_boundary_skipin, sf->skipin_jmin, sf->skipin_kmin, sf->skipin_jmax, sf->skipin_kmax },
            { sf->box_boundary_skipj0, sf->skipj0_imin, sf->skipj0_kmin, sf->skipj0_imax, sf->skipj0_kmax },
            { sf->box_boundary_skipjn, sf->skipjn_imin, sf->skipjn_kmin, sf->skipjn_imax, sf->skipjn_kmax },
            { sf->box_boundary_skipk0, sf->skipk0_imin, sf->skipk0_jmin, sf->skipk0_imax, sf->skipk0_jmax },
            { sf->box_boundary_skipkn, sf->skipkn_imin, sf->skipkn_jmin, sf->skipkn_imax, sf->skipkn_jmax } };
        memcpy(F.skip, sk, sizeof(sk));
        F.nrs = sf->nrs;
        F.ndata = mp->farfield->ndata;
        std::vector<int> ist(F.nrs);
        for (int n = 0; n < F.nrs; n++)
            ist[n] = mp->farfield->rpoints[n].istart;
        int *p;
        if (dupload(E, &p, sf->ri, F.nrs)) return 1;
        F.ri = p;
        if (dupload(E, &p, sf->rj, F.nrs)) return 1;
        F.rj = p;
        if (dupload(E, &p, sf->rk, F.nrs)) return 1;
        F.rk = p;
        if (dupload(E, &p, ist.data(), F.nrs)) return 1;
        F.istart = p;
        size_t cnt = (size_t)F.nrs * 3 * F.ndata;
        if (dmalloc(E, &F.bins, cnt))
            return 1;
        CK(hipMemsetAsync(F.bins, 0, cnt * 8, E->st));
        F.dx = set->sp.dx; F.dy = set->sp.dy; F.dz = set->sp.dz;
        F.dt = set->plan.dt;
        F.c = SV_LIGHT_SPEED;
    }
    if (E->has_dft) {
        GsvDft &D = E->dft;
        std::vector<long long> cells;
        std::vector<int> faces;
        for (int j = F.j0; j < F.j1; j++)
            for (int k = F.k0; k < F.k1; k++) {
                cells.push_back(cell_index(g, F.i0, j, k)); faces.push_back(0);
                cells.push_back(cell_index(g, F.i1, j, k)); faces.push_back(1);
            }
        for (int i = F.i0; i < F.i1; i++)
            for (int k = F.k0; k < F.k1; k++) {
                cells.push_back(cell_index(g, i, F.j0, k)); faces.push_back(2);
                cells.push_back(cell_index(g, i, F.j1, k)); faces.push_back(3);
            }
        for (int i = F.i0; i < F.i1; i++)
            for (int j = F.j0; j < F.j1; j++) {
                cells.push_back(cell_index(g, i, j, F.k0)); faces.push_back(4);
                cells.push_back(cell_index(g, i, j, F.k1)); faces.push_back(5);
            }
        D.ncells = (long long)cells.size();
        D.nw = sf->ndft;
        D.dt = set->plan.dt;
        long long *pc;
        int *pf;
        double *po;
        if (dupload(E, &pc, cells.data(), cells.size()) || dupload(E, &pf, faces.data(), faces.size())
            || dupload(E, &po, mp->farfield->dft_omega, D.nw))
            return 1;
        D.cell = pc;
        D.face = pf;
        D.omega = po;
        size_t cnt = (size_t)D.ncells * D.nw * 8;
        if (dmalloc(E, &D.acc, cnt))
            return 1;
        CK(hipMemsetAsync(D.acc, 0, cnt * 8, E->st));
        D.T = dft_ring_depth(D.nw);
        if (dmalloc(E, &D.ring, (size_t)D.T * 4 * D.ncells))
            return 1;
        E->dft_first = 0;
        E->dft_pending = 0;
        /* the far-field transform of these accumulators, for NFFF_DFT_PATTERN files */
        mp->farfield->dft_transform = gsv_engine_farfield_dft;
        mp->farfield->dft_ctx = E;
    }
    return 0;
}

/* FLUX_DFT planes: one face-6 cell per (plane, j, i) of the rectangle */
static int
setup_flux(Engine *E)
{
    SvSet *set = E->set;
    SvPool *mp = E->mp;
    const SvSetFlux *x = &set->sx;
    memset(&E->flx, 0, sizeof(E->flx));
    E->has_flux = x->nplanes > 0;
    if (!E->has_flux)
        return 0;
    GsvDft &D = E->flx;
    std::vector<long long> cells;
    for (int p = 0; p < x->nplanes; p++)
        for (int j = x->j0; j < x->j1; j++)
            for (int i = x->i0; i < x->i1; i++)
                cells.push_back(cell_index(E->g, i, j, x->k[p]));
    std::vector<int> faces(cells.size(), 6);
    D.ncells = (long long)cells.size();
    D.nw = mp->farfield->nflux;
    D.dt = set->plan.dt;
    long long *pc;
    int *pf;
    double *po;
    if (dupload(E, &pc, cells.data(), cells.size()) || dupload(E, &pf, faces.data(), faces.size())
        || dupload(E, &po, mp->farfield->flux_omega, D.nw))
        return 1;
    D.cell = pc;
    D.face = pf;
    D.omega = po;
    size_t cnt = (size_t)D.ncells * D.nw * 8;
    if (dmalloc(E, &D.acc, cnt))
        return 1;
    CK(hipMemsetAsync(D.acc, 0, cnt * 8, E->st));
    D.T = dft_ring_depth(D.nw);
    if (dmalloc(E, &D.ring, (size_t)D.T * 4 * D.ncells))
        return 1;
    E->flx_first = 0;
    E->flx_pending = 0;
    return 0;
}

/* fold the pending samples of one running DFT into its accumulators */
static int
ring_flush(Engine *E, int on, const GsvDft &D, int &first, int &pending)
{
    if (on && pending > 0) {
        CK(gsv_launch_dft_flush(&D, first, pending, E->st));
        first += pending;
        pending = 0;
    }
    return 0;
}

/* record this step's samples (folding first if the ring is not contiguous with them) */
static int
ring_sample(Engine *E, const GsvDft &D, int &first, int &pending, int step)
{
    if (pending > 0 && step != first + pending && ring_flush(E, 1, D, first, pending))
        return 1;
    if (pending == 0)
        first = step;
    CK(gsv_launch_dft_sample(&E->g, &D, step % D.T, E->st));
    if (++pending == D.T && ring_flush(E, 1, D, first, pending))
        return 1;
    return 0;
}

static int
dft_flush(Engine *E)
{
    return ring_flush(E, E->has_dft, E->dft, E->dft_first, E->dft_pending)
           || ring_flush(E, E->has_flux, E->flx, E->flx_first, E->flx_pending);
}

int
gsv_engine_dft_flush(Engine *E)
{
    return dft_flush(E);
}

/* OUT_SUM boxes: the voxel selection (permittivity within EPSILON_0/100 of
 * the requested one, src3d/output.c:282) and the sigma weight of the
 * absorption mode are fixed, so they become one float weight per box cell */
static int
setup_sums(Engine *E)
{
    SvSet *set = E->set;
    for (int n = 0; n < set->so.nsums; n++) {
        const SvOutputSum *o = set->so.sums + n;
        GsvBoxSum b;
        memset(&b, 0, sizeof(b));
        b.i0 = o->box_i0; b.j0 = o->box_j0; b.k0 = o->box_k0;
        b.ni = o->box_in > o->box_i0 ? o->box_in - o->box_i0 : 0;
        b.nj = o->box_jn > o->box_j0 ? o->box_jn - o->box_j0 : 0;
        b.nk = o->box_kn > o->box_k0 ? o->box_kn - o->box_k0 : 0;
        b.mode = o->component;
        size_t cnt = (size_t)b.ni * b.nj * b.nk;
        if (cnt && (o->layered_epsilon != -1 || o->component == SV_SUM_ABS)) {
            std::vector<float> w(cnt);
            #pragma omp parallel for schedule(static)
            for (int k = 0; k < b.nk; k++)
                for (int j = 0; j < b.nj; j++)
                    for (int i = 0; i < b.ni; i++) {
                        double eps, sig;
                        sv_pool_epsilon_sigma(E->mp, i + b.i0, j + b.j0, k + b.k0, &eps, &sig);
                        int keep = o->layered_epsilon == -1 || fabs(o->layered_epsilon - eps) < SV_EPSILON_0 / 100.0;
                        w[(size_t)i + (size_t)b.ni * (j + (size_t)b.nj * k)] =
                            keep ? (o->component == SV_SUM_ABS ? (float)sig : 1.f) : 0.f;
                    }
            float *dw;
            if (dupload(E, &dw, w.data(), cnt))
                return 1;
            b.w = dw;
        }
        if (dmalloc(E, &b.out, (size_t)set->sc.nsteps + 1))
            return 1;
        CK(hipMemsetAsync(b.out, 0, ((size_t)set->sc.nsteps + 1) * 8, E->st));
        E->sums.push_back(b);
    }
    return 0;
}

extern "C" int
sv_pool_gpu_init(SvPool *mp, SvSet *set)
{
    Engine *E = (Engine *)mp->gpu;
    (void)set;
    if (!E) {
        fprintf(stderr, "Error: sv_pool_gpu_init called without a GPU plan\n");
        return 1;
    }
    GsvGrid &g = E->g;
    CK(hipSetDevice(E->dev));
    CK(hipStreamCreateWithFlags(&E->st, hipStreamNonBlocking));
    E->field_floats = (size_t)g.plane * (g.nz + 2);
    float **fp[6] = { &g.ex, &g.ey, &g.ez, &g.hx, &g.hy, &g.hz };
    for (int q = 0; q < 6; q++) {
        if (dmalloc(E, &E->fbase[q], E->field_floats))
            return 1;
        CK(hipMemsetAsync(E->fbase[q], 0, E->field_floats * 4, E->st));
        *fp[q] = E->fbase[q] + g.plane;
    }
    g.exg = E->fbase[0]; g.eyg = E->fbase[1]; g.ezg = E->fbase[2];
    g.hxg = E->fbase[3]; g.hyg = E->fbase[4]; g.hzg = E->fbase[5];
    {
        uint8_t *d;
        if (dupload(E, &d, E->ecls_host.data(), E->ecls_host.size())) return 1;
        g.ecls = d;
        if (dupload(E, &d, E->hcls_host.data(), E->hcls_host.size())) return 1;
        g.hcls = d;
        GsvECoef *et;
        GsvHCoef *ht;
        if (dupload(E, &et, E->etab.data(), E->etab.size())) return 1;
        if (dupload(E, &ht, E->htab.data(), E->htab.size())) return 1;
        g.etab = et;
        g.htab = ht;
        g.netab = (int)E->etab.size();
        g.nhtab = (int)E->htab.size();
        if (!E->pctab_host.empty()) {
            GsvECoef *pt;
            if (dupload(E, &pt, E->pctab_host.data(), E->pctab_host.size())) return 1;
            g.pctab = pt;
        }
    }
    if (g.rc) {
        GsvRc *rm;
        if (dupload(E, &rm, E->rcm_host, GSV_MAX_ADE))
            return 1;
        g.rcm = rm;
    }
    if (g.ade) {
        GsvAde *am;
        if (dupload(E, &am, E->adem_host, GSV_MAX_ADE))
            return 1;
        g.adem = am;
    }
    if (g.ade) {
        g.pxa = (g.nx + 63) / 64 * 64;
        g.ade_plane_d = (long long)g.ny * g.pxa * GSV_ADE_D;
        size_t n = (size_t)g.ade_plane_d * g.nz;
        if (dmalloc(E, &g.ade_d, n))
            return 1;
        CK(hipMemsetAsync(g.ade_d, 0, n * 4, E->st));
        if (g.ade_poles) {
            g.ade_plane_p = (long long)g.ny * g.pxa * GSV_ADE_P(g.ade_poles);
            n = (size_t)g.ade_plane_p * g.nz;
            if (dmalloc(E, &g.ade_p, n))
                return 1;
            CK(hipMemsetAsync(g.ade_p, 0, n * 4, E->st));
        }
    }
    if (setup_cpml(E) || setup_tsf(E) || setup_tsff(E) || setup_ext(E) || setup_sf(E) || setup_sources_and_outputs(E) || setup_liao_periodic(E) || setup_farfield(E) || setup_pfarfield(E) || setup_ltsf(E) || setup_forces(E)
        || setup_flux(E) || setup_sums(E))
        return 1;
    for (int q = 0; q < 4; q++)
        CK(hipEventCreate(&E->ev[q]));
    CK(hipStreamSynchronize(E->st));
    /* the host copy of the class ids is not needed any more */
    std::vector<uint8_t>().swap(E->ecls_host);
    std::vector<uint8_t>().swap(E->hcls_host);
    E->ready = 1;
    return 0;
}

/* ------------------------------------------------------------------ */
/* host <-> device                                                      */
/* ------------------------------------------------------------------ */

static int
upload_field(Engine *E, const SvDCube *c, float *dst)
{
    const GsvGrid &g = E->g;
    size_t n = (size_t)g.plane * g.nz, q;
    int nonzero = 0;
    for (q = 0; q < (size_t)c->xres * c->yres * c->zres; q++)
        if (c->block[q] != 0) {
            nonzero = 1;
            break;
        }
    if (!nonzero) {
        CK(hipMemsetAsync(dst, 0, n * 4, E->st));
        return 0;
    }
    std::vector<float> h(n, 0.f);
    #pragma omp parallel for schedule(static)
    for (int k = 0; k < g.nz; k++)
        for (int j = 0; j < g.ny; j++)
            for (int i = 0; i < g.nx; i++)
                h[(size_t)i + (size_t)g.px * j + (size_t)g.plane * k] = (float)c->data[i][j][k];
    CK(hipMemcpyAsync(dst, h.data(), n * 4, hipMemcpyHostToDevice, E->st));
    CK(hipStreamSynchronize(E->st));
    return 0;
}

static int
download_field(Engine *E, const float *src, SvDCube *c)
{
    const GsvGrid &g = E->g;
    size_t n = (size_t)g.plane * g.nz;
    std::vector<float> h(n);
    CK(hipMemcpyAsync(h.data(), src, n * 4, hipMemcpyDeviceToHost, E->st));
    CK(hipStreamSynchronize(E->st));
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < g.nx; i++)
        for (int j = 0; j < g.ny; j++)
            for (int k = 0; k < g.nz; k++)
                c->data[i][j][k] = h[(size_t)i + (size_t)g.px * j + (size_t)g.plane * k];
    return 0;
}

extern "C" int
sv_pool_gpu_copyto(SvPool *mp, SvSet *set, SvGpuSync what)
{
    Engine *E = (Engine *)mp->gpu;
    (void)set;
    if (!E || !E->ready)
        return 1;
    CK(hipSetDevice(E->dev));
    SvYeeData *d = mp->d;
    if (what == SV_SYNC_ALL || what == SV_SYNC_EH || what == SV_SYNC_E) {
        if (upload_field(E, d->ex, E->g.ex) || upload_field(E, d->ey, E->g.ey) || upload_field(E, d->ez, E->g.ez))
            return 1;
    }
    if (what == SV_SYNC_ALL || what == SV_SYNC_EH || what == SV_SYNC_H) {
        if (upload_field(E, d->hx, E->g.hx) || upload_field(E, d->hy, E->g.hy) || upload_field(E, d->hz, E->g.hz))
            return 1;
    }
    CK(hipStreamSynchronize(E->st));
    return 0;
}

static int
fetch_aux(Engine *E, SvPool *mp, SvSet *set)
{
    /* traces and far-field bins, src3d/gpu.c:727-780 */
    if (E->npts) {
        int nst = set->sc.nsteps + 1;
        std::vector<float> rec((size_t)nst * E->npts * 6);
        CK(hipMemcpyAsync(rec.data(), E->d_rec, rec.size() * 4, hipMemcpyDeviceToHost, E->st));
        CK(hipStreamSynchronize(E->st));
        for (int n = 0; n < E->npts; n++)
            for (int s = 0; s < nst; s++)
                for (int q = 0; q < 6; q++)
                    mp->out->outpointdata[n][(size_t)s * 6 + q] = rec[((size_t)s * E->npts + n) * 6 + q];
    }
    if (E->has_ff) {
        int nd = E->ff.ndata;
        std::vector<double> b((size_t)E->ff.nrs * 3 * nd);
        CK(hipMemcpyAsync(b.data(), E->ff.bins, b.size() * 8, hipMemcpyDeviceToHost, E->st));
        CK(hipStreamSynchronize(E->st));
        for (int n = 0; n < E->ff.nrs; n++)
            memcpy(mp->farfield->rpoints[n].ex, b.data() + (size_t)n * 3 * nd, 3 * (size_t)nd * 8);
    }
    if (E->has_pff) {
        int nd = E->pff.ndata;
        std::vector<double> b((size_t)E->pff.nrs * 3 * nd);
        CK(hipMemcpyAsync(b.data(), E->pff.bins, b.size() * 8, hipMemcpyDeviceToHost, E->st));
        CK(hipStreamSynchronize(E->st));
        for (int n = 0; n < E->pff.nrs; n++)
            memcpy(mp->pfarfield->prpoints[n].ex, b.data() + (size_t)n * 3 * nd, 3 * (size_t)nd * 8);
    }
    if (E->has_dft || E->has_flux) {
        if (dft_flush(E))
            return 1;
        if (E->has_dft)
            CK(hipMemcpyAsync(mp->farfield->dft_data, E->dft.acc, (size_t)E->dft.ncells * E->dft.nw * 64,
                              hipMemcpyDeviceToHost, E->st));
        if (E->has_flux)
            CK(hipMemcpyAsync(mp->farfield->flux_data, E->flx.acc, (size_t)E->flx.ncells * E->flx.nw * 64,
                              hipMemcpyDeviceToHost, E->st));
        CK(hipStreamSynchronize(E->st));
    }
    for (size_t n = 0; n < E->forces.size(); n++)
        CK(hipMemcpyAsync(mp->out->outforcedata[n], E->forces[n].out, 3 * ((size_t)set->sc.nsteps + 1) * 8,
                          hipMemcpyDeviceToHost, E->st));
    for (size_t n = 0; n < E->sums.size(); n++)
        CK(hipMemcpyAsync(mp->out->outsumdata[n], E->sums[n].out, ((size_t)set->sc.nsteps + 1) * 8,
                          hipMemcpyDeviceToHost, E->st));
    /* the force and sum copies land in pageable host memory the caller reads next */
    CK(hipStreamSynchronize(E->st));
    return 0;
}

extern "C" int
sv_pool_gpu_copyfrom(SvPool *mp, SvSet *set, SvGpuSync what)
{
    Engine *E = (Engine *)mp->gpu;
    if (!E || !E->ready)
        return 1;
    CK(hipSetDevice(E->dev));
    SvYeeData *d = mp->d;
    if (what == SV_SYNC_ALL || what == SV_SYNC_EH || what == SV_SYNC_E) {
        if (download_field(E, E->g.ex, d->ex) || download_field(E, E->g.ey, d->ey) || download_field(E, E->g.ez, d->ez))
            return 1;
    }
    if (what == SV_SYNC_ALL || what == SV_SYNC_EH || what == SV_SYNC_H) {
        if (download_field(E, E->g.hx, d->hx) || download_field(E, E->g.hy, d->hy) || download_field(E, E->g.hz, d->hz))
            return 1;
    }
    return fetch_aux(E, mp, set);
}

/* ------------------------------------------------------------------ */
/* step entry points                                                    */
/* ------------------------------------------------------------------ */

#define ENGINE_OR_FAIL(mp)                                                          \
    Engine *E = (Engine *)(mp)->gpu;                                                \
    if (!E || !E->ready) {                                                          \
        fprintf(stderr, "Error: GPU engine not initialised (%s)\n", __func__);     \
        return 1;                                                                   \
    }

extern "C" int
sv_pool_gpu_boundary_copy(SvPool *mp, SvSet *set)
{
    ENGINE_OR_FAIL(mp);
    (void)set;
    if (E->has_liao)
        CK(gsv_launch_liao_store(&E->g, &E->liao, E->st));
    return 0;
}

extern "C" int
sv_pool_gpu_ystep_h(SvPool *mp, SvSet *set)
{
    ENGINE_OR_FAIL(mp);
    (void)set;
    if (E->timing)
        CK(hipEventRecord(E->ev[0], E->st));
    CK(gsv_launch_hstep(&E->g, E->idbytes_h, E->bx_h, E->by_h, E->kchunk_h, E->st));
    if (E->timing) {
        float ms;
        CK(hipEventRecord(E->ev[1], E->st));
        CK(hipEventSynchronize(E->ev[1]));
        CK(hipEventElapsedTime(&ms, E->ev[0], E->ev[1]));
        E->t_h_ms += ms;
        E->n_h++;
    }
    return 0;
}

extern "C" int
sv_pool_gpu_apply_source_hstep(SvPool *mp, SvSet *set)
{
    ENGINE_OR_FAIL(mp);
    if (E->has_src)
        CK(gsv_launch_point_src(&E->g, &E->src, set->sc.step_act, 1, E->st));
    if (E->has_tsf) {
        CK(gsv_launch_tsf_line(&E->tsf, set->sc.step_act, E->st));
        CK(gsv_launch_tsf_faces(&E->g, &E->tsf, 1, E->st));
    }
    if (E->has_ltsf)
        CK(gsv_launch_ltsf_faces(&E->g, &E->ltsf, 1, set->sc.step_act, E->st));
    if (E->has_tsff) {
        CK(gsv_launch_tsff_line(&E->tsff, set->sc.step_act, E->st));
        CK(gsv_launch_tsff_faces(&E->g, &E->tsff, 1, E->st));
    }
    if (E->has_ext && run_ext(E, 0, 1))
        return 1;
    return 0;
}

extern "C" int
sv_pool_gpu_boundary_hstep(SvPool *mp, SvSet *set)
{
    ENGINE_OR_FAIL(mp);
    (void)set;
    if (E->has_per)
        CK(gsv_launch_periodic(&E->g, &E->per, 1, E->st));
    return 0;
}

extern "C" int
sv_pool_gpu_ystep_e(SvPool *mp, SvSet *set)
{
    ENGINE_OR_FAIL(mp);
    if (E->adj && gsv_adjoint_pre_e(E, set->sc.step_act))
        return 1;
    if (E->timing)
        CK(hipEventRecord(E->ev[2], E->st));
    CK(gsv_launch_estep(&E->g, E->idbytes_e, E->bx_e, E->by_e, E->kchunk_e, E->st));
    if (E->timing) {
        float ms;
        CK(hipEventRecord(E->ev[3], E->st));
        CK(hipEventSynchronize(E->ev[3]));
        CK(hipEventElapsedTime(&ms, E->ev[2], E->ev[3]));
        E->t_e_ms += ms;
        E->n_e++;
    }
    /* the scattered-field source's values on the PEC rule's components (inside
     * ystep_e in the reference, before its boundary E step) */
    if (E->has_sf)
        CK(gsv_launch_sf_pec(&E->g, &E->sfl, E->sf_cells, E->n_sf, E->st));
    /* the rest of the CPU boundary E step, in its order (src3d/boundary.c:521-936) */
    if (E->has_liao) {
        CK(gsv_launch_liao_faces(&E->g, &E->liao, E->st));
        for (int axis = 0; axis < 3; axis++)
            if (E->rezero[axis])
                CK(gsv_launch_zero_planes(&E->g, axis, E->st));
    }
    if (E->has_per)
        CK(gsv_launch_periodic(&E->g, &E->per, 0, E->st));
    if (E->adj && gsv_adjoint_post_e(E, set->sc.step_act))
        return 1;
    return 0;
}

extern "C" int
sv_pool_gpu_apply_source_estep(SvPool *mp, SvSet *set)
{
    ENGINE_OR_FAIL(mp);
    if (E->has_src)
        CK(gsv_launch_point_src(&E->g, &E->src, set->sc.step_act, 0, E->st));
    if (E->has_tsf)
        CK(gsv_launch_tsf_faces(&E->g, &E->tsf, 0, E->st));
    if (E->has_ltsf)
        CK(gsv_launch_ltsf_faces(&E->g, &E->ltsf, 0, set->sc.step_act, E->st));
    if (E->has_tsff)
        CK(gsv_launch_tsff_faces(&E->g, &E->tsff, 0, E->st));
    /* the scattered-field source's line, one step (two at step 0),
     * src3d/source.c:1319-1338 */
    if (E->has_sf) {
        CK(gsv_launch_tsf_line(&E->sfl, set->sc.step_act, E->st));
        if (set->sc.step_act == 0)
            CK(gsv_launch_tsf_line(&E->sfl, set->sc.step_act, E->st));
    }
    if (E->has_ext && run_ext(E, 1, 0))
        return 1;
    return 0;
}

extern "C" int
sv_pool_gpu_boundary_estep(SvPool *mp, SvSet *set)
{
    ENGINE_OR_FAIL(mp);
    (void)E;
    (void)set;
    return 0;
}

extern "C" int
sv_pool_gpu_getsum(SvPool *mp, SvSet *set)
{
    ENGINE_OR_FAIL(mp);
    if (set->sc.step_act > set->sc.nsteps)
        return 1;
    for (size_t n = 0; n < E->sums.size(); n++)
        CK(gsv_launch_box_sum(&E->g, &E->sums[n], set->sc.step_act, E->st));
    return 0;
}

extern "C" int
sv_pool_gpu_apply_farfield(SvPool *mp, SvSet *set)
{
    ENGINE_OR_FAIL(mp);
    if (E->adj && gsv_adjoint_after_step(E, set->sc.step_act))
        return 1;
    if (E->has_ff)
        CK(gsv_launch_farfield_td(&E->g, &E->ff, set->sc.step_act, E->st));
    if (E->has_dft && ring_sample(E, E->dft, E->dft_first, E->dft_pending, set->sc.step_act))
        return 1;
    if (E->has_flux && ring_sample(E, E->flx, E->flx_first, E->flx_pending, set->sc.step_act))
        return 1;
    return 0;
}

extern "C" int
sv_pool_gpu_apply_pfarfield(SvPool *mp, SvSet *set)
{
    ENGINE_OR_FAIL(mp);
    if (E->has_pff)
        CK(gsv_launch_pfarfield(&E->g, &E->pff, set->sc.step_act, E->st));
    return 0;
}

extern "C" int
sv_pool_gpu_getpoints(SvPool *mp, SvSet *set)
{
    ENGINE_OR_FAIL(mp);
    if (E->npts) {
        if (set->sc.step_act > set->sc.nsteps)
            return 1;
        CK(gsv_launch_record(&E->g, E->d_ppos, E->npts,
                             E->d_rec + (size_t)set->sc.step_act * E->npts * 6, E->st));
    }
    return 0;
}

extern "C" int
sv_pool_gpu_getforce(SvPool *mp, SvSet *set)
{
    ENGINE_OR_FAIL(mp);
    for (size_t n = 0; n < E->forces.size(); n++)
        CK(gsv_launch_force(&E->g, &E->forces[n], set->sc.step_act, E->st));
    return 0;
}

/* ------------------------------------------------------------------ */
/* engine helpers used by the session API                               */
/* ------------------------------------------------------------------ */

extern "C" int
gsv_engine_sync(void *eng)
{
    Engine *E = (Engine *)eng;
    CK(hipSetDevice(E->dev));
    CK(hipStreamSynchronize(E->st));
    return 0;
}

/* launch shape of the volume kernels: bx * by <= 512 threads, z chunk >= 1 */
extern "C" int
gsv_engine_set_launch(void *eng, const int *shape)
{
    Engine *E = (Engine *)eng;
    for (int q = 0; q < 2; q++)
        if (shape[3 * q] <= 0 || shape[3 * q + 1] <= 0 || shape[3 * q] * shape[3 * q + 1] > 512 || shape[3 * q + 2] <= 0)
            return 1;
    E->bx_h = shape[0];
    E->by_h = shape[1];
    E->kchunk_h = shape[2];
    E->bx_e = shape[3];
    E->by_e = shape[4];
    E->kchunk_e = shape[5];
    return 0;
}

extern "C" void
gsv_engine_get_launch(void *eng, int *shape)
{
    Engine *E = (Engine *)eng;
    shape[0] = E->bx_h;
    shape[1] = E->by_h;
    shape[2] = E->kchunk_h;
    shape[3] = E->bx_e;
    shape[4] = E->by_e;
    shape[5] = E->kchunk_e;
}

extern "C" void
gsv_engine_set_timing(void *eng, int on)
{
    Engine *E = (Engine *)eng;
    E->timing = on;
    E->t_h_ms = E->t_e_ms = 0;
    E->n_h = E->n_e = 0;
}

/* out[0..3] = total ms of H kernels, count, total ms of E kernels, count */
extern "C" void
gsv_engine_kernel_times(void *eng, double *out)
{
    Engine *E = (Engine *)eng;
    out[0] = E->t_h_ms;
    out[1] = (double)E->n_h;
    out[2] = E->t_e_ms;
    out[3] = (double)E->n_e;
}

/* info[0..7] = E classes, E id bytes, H classes, H id bytes, ADE, poles, bx*by, kchunk_e */
extern "C" void
gsv_engine_info(void *eng, int *info)
{
    Engine *E = (Engine *)eng;
    info[0] = (int)E->etab.size();
    info[1] = E->idbytes_e;
    info[2] = (int)E->htab.size();
    info[3] = E->idbytes_h;
    info[4] = E->g.ade;
    info[5] = E->g.ade_poles;
    info[6] = E->bx_e * E->by_e;
    info[7] = E->kchunk_e;
}

extern "C" int
gsv_engine_get_field(void *eng, int comp, float *out)
{
    Engine *E = (Engine *)eng;
    const GsvGrid &g = E->g;
    float *src[6] = { g.ex, g.ey, g.ez, g.hx, g.hy, g.hz };
    if (comp < 0 || comp > 5)
        return 1;
    size_t n = (size_t)g.plane * g.nz;
    std::vector<float> h(n);
    CK(hipSetDevice(E->dev));
    CK(hipMemcpyAsync(h.data(), src[comp], n * 4, hipMemcpyDeviceToHost, E->st));
    CK(hipStreamSynchronize(E->st));
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < g.nx; i++)
        for (int j = 0; j < g.ny; j++)
            for (int k = 0; k < g.nz; k++)
                out[((size_t)i * g.ny + j) * g.nz + k] = h[(size_t)i + (size_t)g.px * j + (size_t)g.plane * k];
    return 0;
}

extern "C" int
gsv_engine_fetch_aux(void *eng)
{
    Engine *E = (Engine *)eng;
    CK(hipSetDevice(E->dev));
    return fetch_aux(E, E->mp, E->set);
}

extern "C" void
gsv_engine_free(void *eng)
{
    Engine *E = (Engine *)eng;
    if (!E)
        return;
    hipSetDevice(E->dev);
    if (E->st)
        hipStreamSynchronize(E->st);
    gsv_adjoint_free(E);
    for (void *p : E->allocs)
        hipFree(p);
    if (E->ready)
        for (int q = 0; q < 4; q++)
            hipEventDestroy(E->ev[q]);
    if (E->st)
        hipStreamDestroy(E->st);
    delete E;
}
