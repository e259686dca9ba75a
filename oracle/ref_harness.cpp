// ref_harness.cpp -- TEST INFRASTRUCTURE ONLY.
//
// Exposes the parts of the rouckas/mag2d sources that compile on their own
// (Field2D.hpp/.cpp, Field3D.hpp, random.cpp, tabulate.cpp, mymath.cpp,
// histogram.cpp) through extern "C" entry points, so the tests can check the
// oracle's restatement against the reference's own code.  The reference
// sources are compiled where they lie (include path set by oracle/Makefile);
// nothing here stands in for a missing header.  Output goes to oracle/_ref/.
#include "Field2D.hpp"
#include "Field2D.cpp"
#include "Field3D.hpp"
#include "random.cpp"
#include "tabulate.cpp"
#include "mymath.cpp"
#include "histogram.cpp"
#include <cstring>
#include <vector>

extern "C" {

// Field2D::accumulate for n particles into a fresh grid (Field2D.hpp:45-62).
// Returns the number of deposits that threw.
int ref_f2_accumulate_many(int nx, int nz, double dx, double dz, double xmin, double zmin,
                           double q, long n, const double *xs, const double *zs, double *out)
{
    Field2D f(nx, nz, dx, dz, xmin, zmin);
    f.reset();
    int bad = 0;
    for (long k = 0; k < n; k++) {
        try { f.accumulate(q, xs[k], zs[k]); } catch (...) { bad++; }
    }
    std::memcpy(out, f.data[0], sizeof(double) * nx * nz);
    return bad;
}

int ref_f2_interpolate_many(int nx, int nz, double dx, double dz, double xmin, double zmin,
                            const double *data, long n, const double *xs, const double *zs,
                            double *out)
{
    Field2D f(nx, nz, dx, dz, xmin, zmin);
    std::memcpy(f.data[0], data, sizeof(double) * nx * nz);
    int bad = 0;
    for (long k = 0; k < n; k++) {
        try { out[k] = f.interpolate(xs[k], zs[k]); } catch (...) { out[k] = 0; bad++; }
    }
    return bad;
}

// Field2D::grad (Field2D.hpp:80-168); gx/gz keep their input value where the
// reference leaves them untouched.
void ref_f2_grad_many(int nx, int nz, double dx, double dz, double xmin, double zmin,
                      const double *data, long n, const double *xs, const double *zs,
                      double *gx, double *gz)
{
    Field2D f(nx, nz, dx, dz, xmin, zmin);
    std::memcpy(f.data[0], data, sizeof(double) * nx * nz);
    for (long k = 0; k < n; k++) f.grad(xs[k], zs[k], gx[k], gz[k]);
}

int ref_f3_accumulate_many(int nx, int ny, int nz, double dx, double dy, double dz,
                           double q, long n, const double *xs, const double *ys,
                           const double *zs, double *out)
{
    Field3D f(nx, ny, nz, dx, dy, dz);
    f.reset();
    int bad = 0;
    for (long k = 0; k < n; k++) {
        try { f.accumulate(q, xs[k], ys[k], zs[k]); } catch (...) { bad++; }
    }
    std::memcpy(out, f.data[0][0], sizeof(double) * nx * ny * nz);
    return bad;
}

int ref_f3_interpolate_many(int nx, int ny, int nz, double dx, double dy, double dz,
                            const double *data, long n, const double *xs, const double *ys,
                            const double *zs, double *out)
{
    Field3D f(nx, ny, nz, dx, dy, dz);
    std::memcpy(f.data[0][0], data, sizeof(double) * nx * ny * nz);
    int bad = 0;
    for (long k = 0; k < n; k++) {
        try { out[k] = f.interpolate(xs[k], ys[k], zs[k]); } catch (...) { out[k] = 0; bad++; }
    }
    return bad;
}

// t_random (random.cpp): one generator per handle
void *ref_rng_new(unsigned int seed) { return new t_random(seed); }
void ref_rng_free(void *r) { delete (t_random *)r; }
void ref_rng_u32_many(void *r, long n, unsigned int *out)
{
    for (long k = 0; k < n; k++) out[k] = ((t_random *)r)->iuni();
}
void ref_rng_uni_many(void *r, long n, double *out)
{
    for (long k = 0; k < n; k++) out[k] = ((t_random *)r)->uni();
}
void ref_rng_rot_len(void *r, double len, double *x, double *y, double *z)
{
    ((t_random *)r)->rot(len, *x, *y, *z);
}
void ref_rng_rot(void *r, double *x, double *y, double *z) { ((t_random *)r)->rot(*x, *y, *z); }
void ref_rng_deflect(void *r, double angle, double *x, double *y, double *z)
{
    ((t_random *)r)->deflect(angle, *x, *y, *z);
}
double ref_rng_radius(void *r) { return ((t_random *)r)->radius(); }
void ref_rng_rnor_many(void *r, long n, double *out)
{
    for (long k = 0; k < n; k++) out[k] = ((t_random *)r)->rnor();
}

// vec_interpolate (tabulate.cpp:89-141)
void ref_vec_interpolate_many(int n, const double *xs, const double *ys, long m,
                              const double *q, double *out)
{
    std::vector<double> X(xs, xs + n), Y(ys, ys + n);
    vec_interpolate v(X, Y);
    for (long k = 0; k < m; k++) out[k] = v(q[k]);
}

// Field2D::print (Field2D.cpp:30-44) of a grid to a file
void ref_f2_print(int nx, int nz, double dx, double dz, double xmin, double zmin, const double *data,
                  double factor, const char *path)
{
    Field2D f(nx, nz, dx, dz, xmin, zmin);
    std::memcpy(f.data[0], data, sizeof(double) * nx * nz);
    f.print(path, factor);
}

// Field2D::load (Field2D.cpp:46-130): meta = {jmax, lmax, dx, dy, xmin, ymin};
// data gets up to cap values; returns 0, or -1 where the reference throws
int ref_f2_load(const char *path, double *meta, double *data, long cap)
{
    Field2D f;
    try { f.load(path); } catch (...) { return -1; }
    meta[0] = f.jmax; meta[1] = f.lmax; meta[2] = f.GetDx(); meta[3] = f.GetDy();
    meta[4] = f.GetXMin(); meta[5] = f.GetYMin();
    long n = (long)f.jmax * f.lmax;
    std::memcpy(data, f.data[0], sizeof(double) * (n < cap ? n : cap));
    return 0;
}

// mod (mymath.cpp:62-67)
double ref_mod(double x, double y) { return mod(x, y); }

// Histogram::add / mean_tot (histogram.cpp:22-35,68-71)
double ref_hist_mean_tot(int nbins, double lo, double hi, long n, const double *vals)
{
    Histogram h(nbins, lo, hi);
    for (long k = 0; k < n; k++) h.add(vals[k]);
    return h.mean_tot();
}

}  // extern "C"
