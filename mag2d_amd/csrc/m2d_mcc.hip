// m2d_mcc.hip -- MCC set-up: BaseSpecies::lifetime_init / svmax_find
// (particles.cpp:142-206) and the Interaction constants (particles.hpp:49-92),
// packed into the device table read by BaseSpecies::scatter's kernel form.
#include "m2d_objects.hpp"
#include "m2d_scatter.hpp"
#include <cmath>
#include <new>

namespace m2d {

// Interaction::sigma_v (particles.hpp:61-70) for table / constant rate
static double host_sigma_v(const McDev &m, const McInter &it, const std::vector<double> &tab, double v)
{
    if (it.type == M2D_COULOMB) return coulomb_sigma(it, 0.5 * it.mu * v * v) * v;
    if (it.tab_len > 0) {
        const double *xs = tab.data() + it.tab_off;
        const double *ys = tab.data() + (tab.size() / 2) + it.tab_off;
        int n = it.tab_len;
        double x = 0.5 * it.mu * v * v / kQE;
        if (x >= xs[n - 1]) return ys[n - 1] * v;
        if (x <= xs[0]) return ys[0] * v;
        int j1 = 0, j2 = n - 1;
        while (j2 - j1 > 1) {
            int j3 = (j1 + j2) / 2;
            if (x < xs[j3]) j2 = j3; else j1 = j3;
        }
        double w = (x - xs[j1]) / (xs[j2] - xs[j1]);
        return (ys[j1] * (1 - w) + ys[j2] * w) * v;
    }
    (void)m;
    return it.rate;
}

int mcc_sync(m2d_mcc *m)
{
    M2D_HIP(hipMemcpyAsync(m->d, &m->h, sizeof(McDev), hipMemcpyHostToDevice, m->ctx->stream));
    M2D_HIP(hipStreamSynchronize(m->ctx->stream));
    return M2D_OK;
}

}  // namespace m2d

using namespace m2d;

extern "C" {

int m2d_mcc_create(m2d_ctx *ctx, const m2d_mcc_desc *d, m2d_mcc **out)
{
    M2D_CHECK(ctx && d && out, M2D_ERR_BAD_ARG, "m2d_mcc_create: NULL argument");
    *out = nullptr;
    M2D_CHECK(d->n_targets >= 1 && d->n_targets <= kMaxTargets, M2D_ERR_BAD_ARG,
              "m2d_mcc_create: %d targets (1..%d)", d->n_targets, kMaxTargets);
    M2D_CHECK(d->n_inter >= 0 && d->n_inter <= kMaxInter, M2D_ERR_BAD_ARG,
              "m2d_mcc_create: %d interactions (0..%d)", d->n_inter, kMaxInter);
    M2D_CHECK(d->mass > 0, M2D_ERR_BAD_ARG, "m2d_mcc_create: primary mass must be > 0");
    M2D_CHECK(d->targets && (d->n_inter == 0 || d->inter), M2D_ERR_BAD_ARG, "m2d_mcc_create: NULL table");
    M2D_CHECK(d->n_tab == 0 || (d->tab_E && d->tab_sigma), M2D_ERR_BAD_ARG, "m2d_mcc_create: NULL cross sections");
    m2d_mcc *m = new (std::nothrow) m2d_mcc();
    M2D_CHECK(m, M2D_ERR_NO_MEMORY, "m2d_mcc_create: host allocation");
    m->ctx = ctx;
    ctx_retain(ctx);
    McDev &h = m->h;
    h.n_targets = d->n_targets;
    h.mass = d->mass;
    // interactions grouped by target, in declaration order (pic.cpp:82-101)
    int cursor = 0;
    std::vector<double> tab(2 * (size_t)d->n_tab);
    for (int k = 0; k < d->n_tab; k++) { tab[k] = d->tab_E[k]; tab[d->n_tab + k] = d->tab_sigma[k]; }
    for (int t = 0; t < d->n_targets; t++) {
        const m2d_target &T = d->targets[t];
        if (!(T.mass > 0)) { m2d_mcc_destroy(m); return set_error(M2D_ERR_BAD_ARG, "target %d: mass must be > 0", t); }
        McTarget &D = h.t[t];
        D.density = T.density;
        D.mass = T.mass;
        D.vth = sqrt(2.0 * kKB * T.temperature / T.mass) * M_SQRT1_2;  // particles.hpp:180,190
        D.first = cursor;
        D.n = 0;
        D.px = T.partner.x; D.pz = T.partner.z;
        D.pvx = T.partner.vx; D.pvy = T.partner.vy; D.pvz = T.partner.vz;
        D.pn = T.partner.x ? T.partner.n : 0;
        for (int k = 0; k < d->n_inter; k++) {
            const m2d_interaction &I = d->inter[k];
            if (I.target != t) continue;
            if (I.type == M2D_COULOMB && !(d->temperature > 0 && d->density > 0 && d->charge != 0 && T.charge != 0)) {
                m2d_mcc_destroy(m);
                return set_error(M2D_ERR_BAD_ARG, "m2d_mcc_create: COULOMB needs the primary's temperature, "
                                 "density and charge and the target's charge");
            }
            if (I.type < 0 || I.type > M2D_SUPERELASTIC) {
                m2d_mcc_destroy(m);
                return set_error(M2D_ERR_BAD_ARG, "m2d_mcc_create: interaction type %d", I.type);
            }
            if (I.tab_len > 0 && (I.tab_offset < 0 || I.tab_offset + I.tab_len > d->n_tab)) {
                m2d_mcc_destroy(m);
                return set_error(M2D_ERR_BAD_ARG, "m2d_mcc_create: cross-section rows out of range");
            }
            McInter &J = h.in[cursor++];
            J.type = I.type;
            J.DE = I.DE * kQE;                         // particles.hpp:73
            J.rate = I.rate;
            J.cutoff = I.cutoff;
            if (I.type == M2D_LANGEVIN) J.rate *= I.cutoff * I.cutoff;  // particles.hpp:81-82
            J.mu = d->mass * T.mass / (d->mass + T.mass);
            J.tab_off = I.tab_len > 0 ? I.tab_offset : 0;
            J.tab_len = I.tab_len > 0 ? I.tab_len : 0;
            J.lambda_D = J.qq = 0.0;
            if (I.type == M2D_COULOMB) {
                // particles.cpp:19-26
                J.lambda_D = sqrt(kEps0 * kKB * d->temperature / (d->density * d->charge * d->charge));
                J.qq = d->charge * T.charge / (4 * M_PI * kEps0);
                J.tab_len = 0;
                h.partner_update = 1;
            }
            D.n++;
        }
    }
    if (cursor != d->n_inter) {
        m2d_mcc_destroy(m);
        return set_error(M2D_ERR_BAD_ARG, "m2d_mcc_create: interaction with a bad target index");
    }
    // lifetime_init (particles.cpp:142-170): svmax over v in [0, veV(E_max)) with 1000 samples
    double vmax = sqrt(d->E_max * kQE / d->mass * 2.0);
    double rate = 0;
    m->rates.assign(d->n_targets, 0.0);
    for (int t = 0; t < d->n_targets; t++) {
        McTarget &D = h.t[t];
        double dv = vmax / 1000, svmax = 0.0;
        for (double v = 0; v < vmax; v += dv) {
            double sv = 0;
            for (int k = 0; k < D.n; k++) sv += host_sigma_v(h, h.in[D.first + k], tab, v);
            if (std::isnan(sv)) continue;
            if (sv > svmax) svmax = sv;
        }
        D.rate = svmax * D.density;
        m->rates[t] = D.rate;
        rate += D.rate;
    }
    h.lifetime = rate > 0.0 ? 1.0 / rate : INFINITY;
    hipError_t e = hipMalloc((void **)&m->d, sizeof(McDev));
    if (e == hipSuccess && !tab.empty()) e = hipMalloc((void **)&m->tab, sizeof(double) * tab.size());
    if (e != hipSuccess) {
        m2d_mcc_destroy(m);
        return set_error(M2D_ERR_NO_MEMORY, "m2d_mcc_create: %s", hipGetErrorString(e));
    }
    if (!tab.empty()) {
        e = hipMemcpy(m->tab, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice);
        if (e != hipSuccess) { m2d_mcc_destroy(m); return set_error(M2D_ERR_HIP, "m2d_mcc_create: upload"); }
        h.tabE = m->tab;
        h.tabS = m->tab + d->n_tab;
    }
    int r = mcc_sync(m);
    if (r) { m2d_mcc_destroy(m); return r; }
    *out = m;
    return M2D_OK;
}

int m2d_mcc_destroy(m2d_mcc *m)
{
    if (!m) return M2D_OK;
    if (m->ctx) (void)hipStreamSynchronize(m->ctx->stream);
    if (m->d) (void)hipFree(m->d);
    if (m->tab) (void)hipFree(m->tab);
    if (m->ctx) ctx_release(m->ctx);
    delete m;
    return M2D_OK;
}

double m2d_mcc_lifetime(const m2d_mcc *m) { return m ? m->h.lifetime : INFINITY; }

int m2d_mcc_rates(const m2d_mcc *m, double *rates)
{
    M2D_CHECK(m && rates, M2D_ERR_BAD_ARG, "m2d_mcc_rates: NULL argument");
    for (size_t k = 0; k < m->rates.size(); k++) rates[k] = m->rates[k];
    return M2D_OK;
}

int m2d_mcc_set_partner(m2d_mcc *m, int32_t target, const m2d_particles2d *p)
{
    M2D_CHECK(m && p, M2D_ERR_BAD_ARG, "m2d_mcc_set_partner: NULL argument");
    M2D_CHECK(target >= 0 && target < m->h.n_targets, M2D_ERR_BAD_ARG, "m2d_mcc_set_partner: target %d", target);
    McTarget &D = m->h.t[target];
    D.px = p->x; D.pz = p->z; D.pvx = p->vx; D.pvy = p->vy; D.pvz = p->vz;
    D.pn = p->x ? p->n : 0;
    return mcc_sync(m);
}

}  // extern "C"
