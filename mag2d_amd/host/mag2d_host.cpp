// mag2d_host.cpp -- the reference class API (mag2d.hpp) over the C ABI, and
// two C entry points that exercise it end to end: a host-only check of the
// slot container, and a small trap run through Pic.
#include "mag2d.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <functional>
#include <iostream>
#include <sstream>

namespace mag2d {

void check(int status)
{
    if (status != M2D_OK) throw std::runtime_error(m2d_last_error());
}

// double2int (util.cpp:22-28): (int)(x + 1/2), which must lie within eps of x
static int double2int(double x, double eps = 1e-2)
{
    const int r = (int)(x + 0.5);
    if (!(std::fabs(r - x) <= eps))
        throw std::runtime_error("double2int() " + std::to_string(x) + " is not integer\n");
    return r;
}

static inline float dead_float()
{
    uint32_t b = M2D_DEAD_BITS;
    float f;
    std::memcpy(&f, &b, 4);
    return f;
}
static inline bool is_dead(float f)
{
    uint32_t b;
    std::memcpy(&b, &f, 4);
    return b == M2D_DEAD_BITS;
}

// ================================================================ Device
Device::Device(int device, uint64_t seed) { check(m2d_ctx_create(device, seed, &ctx_)); }

Device::~Device()
{
    if (staging_) m2d_free(ctx_, staging_);
    m2d_ctx_destroy(ctx_);
}

void Device::synchronize() { check(m2d_ctx_synchronize(ctx_)); }

uint32_t Device::status(bool clear)
{
    uint32_t f = 0;
    check(m2d_ctx_status(ctx_, &f, clear ? 1 : 0));
    return f;
}

double *Device::staging(size_t n)
{
    if (n > staging_n_) {
        if (staging_) check(m2d_free(ctx_, staging_));
        staging_ = nullptr;
        staging_n_ = 0;
        size_t want = std::max<size_t>(n, 4096);
        check(m2d_malloc(ctx_, want * sizeof(double), &staging_));
        staging_n_ = want;
    }
    return (double *)staging_;
}

DeviceBuffer::DeviceBuffer(std::shared_ptr<Device> dev, size_t bytes) : dev_(std::move(dev)), bytes_(bytes)
{
    check(m2d_malloc(dev_->ctx(), std::max<size_t>(bytes, 16), &p_));
}

DeviceBuffer::~DeviceBuffer()
{
    if (p_ && dev_) m2d_free(dev_->ctx(), p_);
}

DeviceBuffer::DeviceBuffer(DeviceBuffer &&o) noexcept : dev_(std::move(o.dev_)), p_(o.p_), bytes_(o.bytes_)
{
    o.p_ = nullptr;
    o.bytes_ = 0;
}

DeviceBuffer &DeviceBuffer::operator=(DeviceBuffer &&o) noexcept
{
    if (this != &o) {
        if (p_ && dev_) m2d_free(dev_->ctx(), p_);
        dev_ = std::move(o.dev_);
        p_ = o.p_;
        bytes_ = o.bytes_;
        o.p_ = nullptr;
        o.bytes_ = 0;
    }
    return *this;
}

// ================================================================ ParticleList
void ParticleList::remove(int n)
{
    if (particles.at(n).empty) throw std::runtime_error("removing empty particle");
    particles[n].empty = true;
    empty_.push_back(n);
}

int ParticleList::insert()
{
    if (empty_.empty()) resize((unsigned int)(particles.size() * 1.1) + 1);
    int tmp = empty_.back();
    if (!particles[tmp].empty) throw std::runtime_error("error inserting particle");
    empty_.pop_back();
    particles[tmp].empty = false;
    return tmp;
}

void ParticleList::resize(unsigned int newsize)
{
    size_t old = particles.size();
    if (newsize <= old) return;
    particles.resize(newsize);
    for (size_t i = old; i < newsize; i++) {
        particles[i] = t_particle{0, 0, 0, 0, 0, 0, 0, true};
        empty_.push_back((int)i);
    }
}

void ParticleList::rebuild_free_list()
{
    empty_.clear();
    for (size_t i = 0; i < particles.size(); i++)
        if (particles[i].empty) empty_.push_back((int)i);
}

void ParticleList::save(const std::string &filename) const
{
    std::ofstream fw(filename.c_str(), std::ios::out | std::ios::binary);
    if (!fw.is_open()) throw std::runtime_error("Species::save(): failed opening file");
    int size = (int)particles.size();
    int live = (int)(particles.size() - empty_.size());
    fw.write((const char *)&size, sizeof(int));
    fw.write((const char *)&live, sizeof(int));
    int written = 0;
    for (const t_particle &p : particles) {
        if (p.empty) continue;
        if (++written > live) break;
        fw.write((const char *)&p, sizeof(t_particle));
    }
    if (!fw.good()) throw std::runtime_error("Species::save(): write error");
}

void ParticleList::load(const std::string &filename)
{
    std::ifstream fr(filename.c_str(), std::ios::in | std::ios::binary);
    if (!fr.is_open()) throw std::runtime_error("Species::load(): failed opening file");
    int size = 0, live = 0;
    fr.read((char *)&size, sizeof(int));
    fr.read((char *)&live, sizeof(int));
    if (!fr.good() || size < 0 || live < 0 || live > size)
        throw std::runtime_error("Species::load(): bad header");
    particles.assign((size_t)size, t_particle{0, 0, 0, 0, 0, 0, 0, true});
    for (int i = 0; i < live; i++) {
        fr.read((char *)&particles[i], sizeof(t_particle));
        if (!fr.good()) throw std::runtime_error("Species::load(): read error");
        particles[i].empty = false;
    }
    empty_.clear();
    for (int i = live; i < size; i++) empty_.push_back(i);
}

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11): the library's draw
// stream (seed, stream, slot, step, purpose) on the host, first block
static void philox_block(uint64_t seed, uint32_t stream, uint32_t slot, uint64_t step, uint32_t purpose,
                         uint32_t out[4])
{
    uint32_t c0 = slot, c1 = (uint32_t)step, c2 = 0, c3 = (purpose << 24) | ((uint32_t)(step >> 32) & 0xFFFFFFu);
    uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32) ^ stream;
    for (int r = 0; r < 10; r++) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0, n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c0 = n0;
        c1 = (uint32_t)p1;
        c2 = n2;
        c3 = (uint32_t)p0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

constexpr uint32_t kPurposeReload = 5;

void reload_remap(std::vector<t_particle> &particles, int live, double x_min, double x_max, double z_min,
                  double z_max, uint64_t seed, uint32_t stream)
{
    // uniform in (0,1): (w + 1/2) 2^-32, as the library draws them
    auto uni = [](uint32_t w) { return ((double)w + 0.5) * 2.3283064365386963e-10; };
    for (int i = 0; i < live && i < (int)particles.size(); i++) {
        t_particle &p = particles[(size_t)i];
        const bool ox = p.x < x_min || p.x > x_max, oz = p.z < z_min || p.z > z_max;
        if (!ox && !oz) continue;
        uint32_t w[4];
        philox_block(seed, stream, (uint32_t)i, 0, kPurposeReload, w);
        if (ox) p.x = (x_max - x_min) * uni(w[0]) + x_min;
        if (oz) p.z = (z_max - z_min) * uni(w[1]) + z_min;
    }
}

// ================================================================ Field2D
Field2D::Field2D(std::shared_ptr<Device> dev, int x_sampl, int y_sampl, double dx, double dy, double xmin,
                 double ymin)
    : jmax(x_sampl), lmax(y_sampl), dev_(std::move(dev))
{
    if (x_sampl < 2 || y_sampl < 2 || !(dx > 0) || !(dy > 0)) throw std::runtime_error("Field2D: bad shape");
    g_ = m2d_grid2d{x_sampl, y_sampl, dx, dy, xmin, ymin};
    const size_t n = (size_t)x_sampl * y_sampl;
    own_ = DeviceBuffer(dev_, n * sizeof(double));
    d_ = (double *)own_.get();
    h_.assign(n, 0.0);
    check(m2d_memset(dev_->ctx(), d_, 0, n * sizeof(double)));
    state_ = SYNCED;
}

Field2D::Field2D(std::shared_ptr<Device> dev, double *device_data, int x_sampl, int y_sampl, double dx,
                 double dy, double xmin, double ymin)
    : jmax(x_sampl), lmax(y_sampl), dev_(std::move(dev)), d_(device_data)
{
    g_ = m2d_grid2d{x_sampl, y_sampl, dx, dy, xmin, ymin};
    h_.assign((size_t)x_sampl * y_sampl, 0.0);
    state_ = DEVICE_NEWER;
}

void Field2D::pull() const
{
    if (state_ == DEVICE_NEWER) {
        check(m2d_memcpy_d2h(dev_->ctx(), h_.data(), d_, h_.size() * sizeof(double)));
        state_ = SYNCED;
    }
}

double *Field2D::operator[](int i)
{
    pull();
    state_ = HOST_NEWER;
    return h_.data() + (size_t)i * lmax;
}

const double *Field2D::operator[](int i) const
{
    pull();
    return h_.data() + (size_t)i * lmax;
}

double *Field2D::device()
{
    if (state_ == HOST_NEWER) upload();
    return d_;
}

void Field2D::download() { pull(); }

void Field2D::upload()
{
    check(m2d_memcpy_h2d(dev_->ctx(), d_, h_.data(), h_.size() * sizeof(double)));
    state_ = SYNCED;
}

void Field2D::reset()
{
    check(m2d_memset(dev_->ctx(), d_, 0, h_.size() * sizeof(double)));
    state_ = DEVICE_NEWER;
}

void Field2D::add(const Field2D &other)
{
    if (other.jmax != jmax || other.lmax != lmax)
        throw std::runtime_error("Array2D.add(): matrix sizes don't match\n");
    double *src = const_cast<Field2D &>(other).device();
    check(m2d_field_add(dev_->ctx(), device(), src, 1.0, h_.size()));
    state_ = DEVICE_NEWER;
}

void Field2D::accumulate(double q, double x, double y) { accumulate(q, &x, &y, 1); }

void Field2D::accumulate(double q, const double *x, const double *y, size_t n)
{
    if (n == 0) return;
    // a store of n particles in the staging buffer: x, z, then three zero
    // velocity arrays (16-byte aligned)
    const size_t stride = (n + 3) & ~(size_t)3;
    float *s = (float *)dev_->staging((5 * stride + 1) / 2 + 2);
    std::vector<float> hs(5 * stride, 0.0f);
    for (size_t k = 0; k < n; k++) {
        hs[k] = (float)x[k];
        hs[stride + k] = (float)y[k];
    }
    check(m2d_memcpy_h2d(dev_->ctx(), s, hs.data(), hs.size() * sizeof(float)));
    m2d_particles2d p{s, s + stride, s + 2 * stride, s + 3 * stride, s + 4 * stride, (int64_t)n};
    (void)dev_->status(true);
    check(m2d_accumulate_2d(dev_->ctx(), &g_, &p, q, device()));
    state_ = DEVICE_NEWER;
    if (dev_->status(true) & M2D_FLAG_OUT_OF_RANGE)
        throw std::runtime_error("Field2D::accumulate: outside of range");
}

void Field2D::interpolate(const double *x, const double *y, size_t n, double *out)
{
    if (n == 0) return;
    double *s = dev_->staging(3 * n);
    check(m2d_memcpy_h2d(dev_->ctx(), s, x, n * sizeof(double)));
    check(m2d_memcpy_h2d(dev_->ctx(), s + n, y, n * sizeof(double)));
    check(m2d_interpolate_2d(dev_->ctx(), &g_, device(), s, s + n, (int64_t)n, s + 2 * n));
    check(m2d_memcpy_d2h(dev_->ctx(), out, s + 2 * n, n * sizeof(double)));
}

double Field2D::interpolate(double x, double y)
{
    double v = 0;
    interpolate(&x, &y, 1, &v);
    return v;
}

void Field2D::grad(const double *x, const double *y, size_t n, double *gx, double *gy)
{
    if (n == 0) return;
    double *s = dev_->staging(4 * n);
    check(m2d_memcpy_h2d(dev_->ctx(), s, x, n * sizeof(double)));
    check(m2d_memcpy_h2d(dev_->ctx(), s + n, y, n * sizeof(double)));
    check(m2d_grad_2d(dev_->ctx(), &g_, device(), s, s + n, (int64_t)n, s + 2 * n, s + 3 * n));
    check(m2d_memcpy_d2h(dev_->ctx(), gx, s + 2 * n, n * sizeof(double)));
    check(m2d_memcpy_d2h(dev_->ctx(), gy, s + 3 * n, n * sizeof(double)));
}

void Field2D::grad(double x, double y, double &gx, double &gy) { grad(&x, &y, 1, &gx, &gy); }

void Field2D::print(std::ostream &out, double factor)
{
    pull();
    field_text_write(out, g_, h_.data(), factor);
}

void Field2D::print(const char *filename, double factor)
{
    std::ofstream out(filename);
    print(out, factor);
}

void Field2D::load(const char *filename)
{
    FieldText t = field_text_read(filename);
    if (t.g.nx != jmax || t.g.nz != lmax) throw std::runtime_error("Field2D::load(): file shape differs from the field");
    // print writes each coordinate to 6 significant digits (relative error
    // 5e-6), so the origin, and the spacing found from the first difference
    // of two coordinates, agree with the field's to 1e-5 of the largest
    // coordinate.  The field keeps its own exact grid (the reference takes
    // the file's rounded spacing; the two agree to that rounding).
    auto span = [](double lo, double h, int n) { return std::fabs(lo) + std::fabs(lo + (n - 1) * h) + h; };
    const double tx = 1e-5 * span(g_.xmin, g_.dx, jmax), tz = 1e-5 * span(g_.zmin, g_.dz, lmax);
    if (std::fabs(t.g.dx - g_.dx) > tx || std::fabs(t.g.dz - g_.dz) > tz || std::fabs(t.g.xmin - g_.xmin) > tx ||
        std::fabs(t.g.zmin - g_.zmin) > tz)
        throw std::runtime_error("Field2D::load(): file grid differs from the field");
    h_ = std::move(t.data);
    state_ = HOST_NEWER;
}

// the text format of Field2D::print (Field2D.cpp:30-44): "x<TAB>z<TAB>value"
// per node in the stream's default notation (6 significant digits), a blank
// line after each x row
void field_text_write(std::ostream &out, const m2d_grid2d &g, const double *data, double factor)
{
    for (int j = 0; j < g.nx; j++) {
        for (int l = 0; l < g.nz; l++)
            out << j * g.dx + g.xmin << "\t" << l * g.dz + g.zmin << "\t" << data[(size_t)j * g.nz + l] * factor
                << std::endl;
        out << std::endl;
    }
}

// Field2D::load (Field2D.cpp:46-130): lines of three numbers in any order of
// nodes; spacing, origin and extent from the first change of each coordinate
// (a negative step means the file runs backwards); every node exactly once
FieldText field_text_read(const char *filename)
{
    std::ifstream fr(filename);
    if (fr.fail()) throw std::runtime_error("Field2D::load(): failed opening file\n");
    std::vector<double> xv, yv, fv;
    std::string line;
    while (std::getline(fr, line)) {
        std::istringstream s(line);
        double a, b, c;
        if (s >> a >> b >> c) {
            xv.push_back(a);
            yv.push_back(b);
            fv.push_back(c);
        }
    }
    if (xv.size() < 2) throw std::runtime_error("Field2D::load(): no data");
    auto axis = [](const std::vector<double> &v, double &d, double &vmin, double &vmax) {
        vmin = v.front();
        vmax = v.back();
        size_t i = 1;
        while (i < v.size() && v[i] - v[i - 1] == 0.0) i++;
        if (i == v.size()) throw std::runtime_error("Field2D::load(): constant coordinate");
        d = v[i] - v[i - 1];
        if (d < 0) {
            d = -d;
            std::swap(vmin, vmax);
        }
    };
    double dx, xmin, xmax, dy, ymin, ymax;
    axis(xv, dx, xmin, xmax);
    axis(yv, dy, ymin, ymax);
    const int xs = double2int((xmax - xmin) / dx + 1, 1e-1), ys = double2int((ymax - ymin) / dy + 1, 1e-1);
    if ((size_t)xs * ys != xv.size())
        throw std::runtime_error("Fields::load_magnetic_field() wrong size of input vector");
    FieldText t;
    t.g = m2d_grid2d{xs, ys, dx, dy, xmin, ymin};
    t.data.assign((size_t)xs * ys, std::nan(""));
    for (size_t i = 0; i < xv.size(); i++) {
        int xi = double2int((xv[i] - xmin) / dx, 1e-1), yi = double2int((yv[i] - ymin) / dy, 1e-1);
        // a node beyond the extent found from the ends (a file whose nodes do
        // not run monotonically): the reference writes outside its array there
        if (xi < 0 || xi >= xs || yi < 0 || yi >= ys)
            throw std::runtime_error("Field2D::load(): node outside the grid");
        t.data[(size_t)xi * ys + yi] = fv[i];
    }
    for (double v : t.data)
        if (std::isnan(v)) throw std::runtime_error("Fields::load_magnetic_field() garbage loaded");
    return t;
}

// ================================================================ Param
void Param::derive()
{
    dx = x_max / (x_sampl - 1);
    dz = z_max / (z_sampl - 1);
    idx = 1.0 / dx;
    idz = 1.0 / dz;
    V = n_particles_total / density_total;
    dV = V / ((double)(x_sampl - 1) * (z_sampl - 1));
}

void Param::check() const
{
    if (selfconsistent && rf) throw std::runtime_error("Param: selfconsistent RF simulation not implemented");
    if (selfconsistent && electric_field_from_file)
        throw std::runtime_error("Param: selfconsistent with electric_field_from_file not implemented");
    if (x_sampl < 3 || z_sampl < 3) throw std::runtime_error("Param: need at least 3 samples per axis");
}

// ================================================================ t_grid
t_grid::t_grid(const Param &p) : jmax(p.x_sampl), lmax(p.z_sampl), dx(p.dx), dz(p.dz)
{
    if (!(p.dx > 0)) throw std::runtime_error("t_grid: Param::derive() was not called");
    mask.assign((size_t)jmax * lmax, 0);
    voltage.assign((size_t)jmax * lmax, 0.0);
    m2d_geom2d geo{p.geometry, p.x_max, p.z_max, p.probe_radius, p.u_probe, p.extern_field};
    m2d_grid2d g = grid();
    check(m2d_grid_build_2d(&g, &geo, mask.data(), voltage.data()));
}

m2d_grid2d t_grid::grid() const { return m2d_grid2d{jmax, lmax, dx, dz, 0.0, 0.0}; }

void t_grid::square_electrode(double rmin, double rmax, double zmin, double zmax, double v, mask_type type)
{
    m2d_grid2d g = grid();
    check(m2d_grid_square_electrode_2d(&g, mask.data(), voltage.data(), rmin, rmax, zmin, zmax, v, type));
}

void t_grid::circle_electrode(double x, double z, double r, double v, mask_type type)
{
    m2d_grid2d g = grid();
    check(m2d_grid_circle_electrode_2d(&g, mask.data(), voltage.data(), x, z, r, v, type));
}

bool t_grid::is_free(double x, double z) const
{
    m2d_grid2d g = grid();
    int r = m2d_grid_is_free_2d(&g, mask.data(), x, z);
    if (r < 0) check(r);
    return r != 0;
}

// ================================================================ Fields
static m2d_fields2d_params fields_params(const Param &p)
{
    m2d_fields2d_params fp{};
    fp.grid = m2d_grid2d{p.x_sampl, p.z_sampl, p.dx, p.dz, p.x_min, p.z_min};
    fp.x_max = p.x_max;
    fp.z_max = p.z_max;
    fp.geometry_empty = p.geometry == EMPTY;
    fp.selfconsistent = p.selfconsistent;
    fp.electric_field_from_file = p.electric_field_from_file;
    fp.boundary = p.boundary;
    fp.extern_field = p.extern_field;
    fp.rf = p.rf;
    fp.rf_amplitude = p.rf_amplitude;
    fp.rf_omega = p.rf_omega;
    fp.rf_U0 = p.rf_U0;
    fp.Br = p.Br;
    fp.Bz = p.Bz;
    fp.Bt = p.Bt;
    fp.dV = p.dV;
    return fp;
}

static m2d_fields2d *create_fields(Param &p, Device &dev, const t_grid &grid)
{
    p.check();
    m2d_fields2d_params fp = fields_params(p);
    m2d_fields2d *f = nullptr;
    check(m2d_fields2d_create(dev.ctx(), &fp, grid.mask.data(), grid.voltage.data(), &f));
    return f;
}

Fields::Fields(Param &p, std::shared_ptr<Device> dev) : Fields(p, dev, t_grid(p)) {}

Fields::Fields(Param &p, std::shared_ptr<Device> dev, const t_grid &g)
    : p_(&p), dev_(dev), f_(create_fields(p, *dev, g)), grid(g),
      u(dev, m2d_fields2d_data(f_, M2D_FIELD_U), p.x_sampl, p.z_sampl, p.dx, p.dz, p.x_min, p.z_min),
      uRF(dev, m2d_fields2d_data(f_, M2D_FIELD_URF), p.x_sampl, p.z_sampl, p.dx, p.dz, p.x_min, p.z_min),
      rho(dev, m2d_fields2d_data(f_, M2D_FIELD_RHO), p.x_sampl, p.z_sampl, p.dx, p.dz, p.x_min, p.z_min)
{
}

Fields::~Fields()
{
    if (f_) m2d_fields2d_destroy(f_);
}

m2d_fields2d *Fields::handle()
{
    // host edits of u / uRF reach the device and the gather tables
    if (u.host_newer() || uRF.host_newer()) update();
    return f_;
}

void Fields::u_smooth(bool symmetry, double radius, Field2D *field)
{
    int which = M2D_FIELD_U;
    if (field == &uRF) which = M2D_FIELD_URF;
    else if (field && field != &u) throw std::runtime_error("Fields::u_smooth: smooths u or uRF");
    (void)handle();  // host edits of u / uRF reach the device first
    check(m2d_fields2d_u_smooth(f_, which, symmetry ? 1 : 0, radius));
    (which == M2D_FIELD_U ? u : uRF).mark_device_written();
}

void Fields::update()
{
    (void)u.device();
    (void)uRF.device();
    check(m2d_fields2d_update(f_));
}

void Fields::run_solve(int mode)
{
    (void)rho.device();
    check(m2d_fields2d_solve(handle(), mode, p_->solve_tolerance, &info_));
    // the rhs is built in place on rho, as the reference does
    rho.mark_device_written();
    (mode == M2D_SOLVE_RF ? uRF : u).mark_device_written();
}

void Fields::boundary_solve() { run_solve(M2D_SOLVE_STATIC); }
void Fields::boundary_solve_rf() { run_solve(M2D_SOLVE_RF); }
void Fields::solve() { run_solve(M2D_SOLVE_RAW); }
void Fields::reset() { rho.reset(); }

void Fields::E(double x, double y, double &gx, double &gy, double t)
{
    double *s = dev_->staging(4);
    m2d_fields2d *f = handle();
    double pos[2] = {x, y}, out[2];
    check(m2d_memcpy_h2d(dev_->ctx(), s, pos, sizeof pos));
    check(m2d_fields2d_E(f, s, s + 1, 1, t, s + 2, s + 3));
    check(m2d_memcpy_d2h(dev_->ctx(), out, s + 2, sizeof out));
    gx = out[0];
    gy = out[1];
}

void Fields::B(double x, double y, double &Br, double &Bz, double &Bt) const
{
    if (p_->magnetic_field_const) {
        // fields.hpp:154-169
        Br = p_->Br;
        Bz = p_->Bz;
        Bt = p_->Bt;
        return;
    }
    // fields.hpp:171-176, through the device tables of the passes
    double *s = dev_->staging(5);
    double pos[2] = {x, y}, out[3];
    check(m2d_memcpy_h2d(dev_->ctx(), s, pos, sizeof pos));
    check(m2d_fields2d_B(f_, s, s + 1, 1, s + 2, s + 3, s + 4));
    check(m2d_memcpy_d2h(dev_->ctx(), out, s + 2, sizeof out));
    Br = out[0];
    Bz = out[1];
    Bt = out[2];
}

void Fields::load_magnetic_field(const char *fname)
{
    std::ifstream fr(fname);
    if (fr.fail()) throw std::runtime_error("Fields::load_magnetic_field(): failed opening file\n");
    std::vector<double> rv, zv, brv, bzv;
    std::string line;
    while (std::getline(fr, line)) {
        std::istringstream ss(line);
        double a, b, c, d;
        if (ss >> a >> b >> c >> d) {
            rv.push_back(a);
            zv.push_back(b);
            brv.push_back(c);
            bzv.push_back(d);
        }
    }
    const size_t n = rv.size();
    // spacing, first and last sample of one axis (fields.cpp:902-928): the
    // first change between consecutive rows; a negative one swaps the ends
    auto axis = [&](const std::vector<double> &v, double &h, double &lo, double &hi) {
        size_t i = 1;
        while (i < n && v[i] - v[i - 1] == 0.0) i++;
        if (i >= n) throw std::runtime_error("Fields::load_magnetic_field() wrong size of input vector");
        h = v[i] - v[i - 1];
        lo = v[0];
        hi = v[n - 1];
        if (h < 0) {
            h = -h;
            std::swap(lo, hi);
        }
        return double2int((hi - lo) / h + 1);
    };
    if (n < 4) throw std::runtime_error("Fields::load_magnetic_field() wrong size of input vector");
    double dr, rmin, rmax, dz, zmin, zmax;
    const int rs = axis(rv, dr, rmin, rmax), zs = axis(zv, dz, zmin, zmax);
    if ((size_t)rs * (size_t)zs != n)
        throw std::runtime_error("Fields::load_magnetic_field() wrong size of input vector");
    std::vector<double> br((size_t)rs * zs, NAN), bz((size_t)rs * zs, NAN);
    for (size_t k = 0; k < n; k++) {
        const int ri = double2int((rv[k] - rmin) / dr), zi = double2int((zv[k] - zmin) / dz);
        if (ri < 0 || ri >= rs || zi < 0 || zi >= zs)
            throw std::runtime_error("Fields::load_magnetic_field() garbage loaded");
        br[(size_t)ri * zs + zi] = brv[k];
        bz[(size_t)ri * zs + zi] = bzv[k];
    }
    // a sample left unset is NaN: the library rejects it ("garbage loaded")
    m2d_grid2d g{rs, zs, dr, dz, rmin, zmin};
    check(m2d_fields2d_set_magnetic_field(f_, &g, br.data(), bz.data()));
}

// ================================================================ Species
Species::Species(const SpeciesParams &sp, Param &param, Fields &f, uint32_t stream)
    : name(sp.name), particle(sp.particle), mass(sp.mass), charge(sp.charge), temperature(sp.temperature),
      density(sp.density), dt(sp.dt), t(0), lifetime(INFINITY),
      E_max(sp.E_max > 0. ? sp.E_max : sp.temperature * physconst::k_B / physconst::q_e * 10.0),
      v_max(std::sqrt(2.0 * physconst::k_B * sp.temperature / sp.mass)),
      rho(f.device(), param.x_sampl, param.z_sampl, param.dx, param.dz), field(&f), p_(&param),
      dev_(f.device()), stream_(stream), params_(sp)
{
    if (!(mass > 0) || !(dt > 0)) throw std::runtime_error("Species " + name + ": mass and dt must be > 0");
}

Species::~Species()
{
    if (mcc_) m2d_mcc_destroy(mcc_);
}

m2d_particles2d Species::store()
{
    return m2d_particles2d{(float *)soa_[0].get(), (float *)soa_[1].get(), (float *)soa_[2].get(),
                           (float *)soa_[3].get(), (float *)soa_[4].get(), cap_};
}

void Species::reserve_device(int64_t n)
{
    if (n == cap_) return;
    for (int c = 0; c < 5; c++) soa_[c] = DeviceBuffer(dev_, (size_t)std::max<int64_t>(n, 1) * sizeof(float));
    cap_ = n;
}

void Species::upload()
{
    const int64_t n = (int64_t)particles.size();
    reserve_device(n);
    std::vector<float> h[5];
    for (auto &v : h) v.resize((size_t)n);
    for (int64_t i = 0; i < n; i++) {
        const t_particle &p = particles[(size_t)i];
        h[0][i] = p.empty ? dead_float() : (float)p.x;
        h[1][i] = (float)p.z;
        h[2][i] = (float)p.vx;
        h[3][i] = (float)p.vy;
        h[4][i] = (float)p.vz;
    }
    for (int c = 0; c < 5; c++)
        if (n) check(m2d_memcpy_h2d(dev_->ctx(), soa_[c].get(), h[c].data(), (size_t)n * sizeof(float)));
    if (multicoll()) {
        ensure_ttd();
        std::vector<double> t((size_t)n);
        for (int64_t i = 0; i < n; i++) t[(size_t)i] = particles[(size_t)i].time_to_death;
        if (n) check(m2d_memcpy_h2d(dev_->ctx(), ttd_.get(), t.data(), (size_t)n * sizeof(double)));
    }
    m2d_particles2d s = store();
    check(m2d_sort_forget_2d(dev_->ctx(), &s));
    host_newer_ = dev_newer_ = false;
}

void Species::download()
{
    if (!dev_newer_) return;
    const int64_t n = cap_;
    std::vector<float> h[5];
    for (int c = 0; c < 5; c++) {
        h[c].resize((size_t)n);
        if (n) check(m2d_memcpy_d2h(dev_->ctx(), h[c].data(), soa_[c].get(), (size_t)n * sizeof(float)));
    }
    // y is not on the device, nor is time_to_death outside the multicoll
    // mover: slots keep them where they persist
    std::vector<double> t;
    if (multicoll() && ttd_.bytes() == (size_t)std::max<int64_t>(n, 1) * sizeof(double)) {
        t.resize((size_t)n);
        if (n) check(m2d_memcpy_d2h(dev_->ctx(), t.data(), ttd_.get(), (size_t)n * sizeof(double)));
    }
    particles.resize((size_t)n, t_particle{0, 0, 0, 0, 0, 0, 0, true});
    for (int64_t i = 0; i < n; i++) {
        t_particle &p = particles[(size_t)i];
        p.empty = is_dead(h[0][i]);
        if (p.empty) continue;
        p.x = h[0][i];
        p.z = h[1][i];
        p.vx = h[2][i];
        p.vy = h[3][i];
        p.vz = h[4][i];
        if (!t.empty()) p.time_to_death = t[(size_t)i];
    }
    rebuild_free_list();
    dev_newer_ = false;
}

void Species::ensure_device()
{
    if (host_newer_) upload();
}

void Species::ensure_host() { download(); }

void Species::load(const std::string &filename)
{
    ParticleList::load(filename);
    reload_remap(particles, (int)(particles.size() - empty_.size()), p_->x_min, p_->x_max, p_->z_min, p_->z_max,
                 p_->seed, stream_);
    // the host list is the store now: the next pass uploads it
    dev_newer_ = false;
    host_newer_ = true;
}

int Species::insert()
{
    ensure_host();
    host_newer_ = true;
    return ParticleList::insert();
}

void Species::remove(int n)
{
    ensure_host();
    host_newer_ = true;
    ParticleList::remove(n);
}

void Species::resize(unsigned int newsize)
{
    ensure_host();
    host_newer_ = true;
    ParticleList::resize(newsize);
}

int Species::n_particles()
{
    if (!dev_newer_) return ParticleList::n_particles();
    m2d_particles2d s = store();
    int64_t live = 0;
    check(m2d_count_live_2d(dev_->ctx(), &s, &live));
    return (int)live;
}

m2d_species2d Species::desc() const
{
    m2d_species2d d{};
    d.charge = charge;
    d.mass = mass;
    d.dt = dt;
    d.lifetime = lifetime;
    d.niter = niter;
    d.stream = stream_;
    return d;
}

// the multicoll mover's time_to_death array, one FP64 per slot; slots that
// had none start at 0 (a value-initialised t_particle), so their first event
// comes at once
void Species::ensure_ttd()
{
    const size_t bytes = (size_t)std::max<int64_t>(cap_, 1) * sizeof(double);
    if (ttd_.bytes() == bytes) return;
    ttd_ = DeviceBuffer(dev_, bytes);
    check(m2d_memset(dev_->ctx(), ttd_.get(), 0, bytes));
}

// point the MCC targets that are live species at their current stores
void Species::bind_partners()
{
    for (size_t k = 0; k < partners_.size(); k++)
        if (partners_[k]) {
            partners_[k]->ensure_device();
            m2d_particles2d ps = partners_[k]->store();
            check(m2d_mcc_set_partner(mcc_, (int32_t)k, &ps));
        }
}

void Species::run_pass(uint32_t flags)
{
    ensure_device();
    if (cap_ == 0) return;
    bind_partners();
    m2d_species2d d = desc();
    m2d_particles2d s = store();
    const bool dep = (flags & M2D_STEP_DEPOSIT) != 0;
    check(m2d_advance_2d(dev_->ctx(), field->handle(), &d, &s, mcc_, dep ? rho.device() : nullptr, flags,
                         nullptr));
    if (dep) rho.mark_device_written();
    dev_newer_ = true;
}

void Species::advance()
{
    // advance_position + advance_boundary (particles.hpp:342-349); the mover
    // of Param::mover (particles.cpp:775-792)
    if (p_->mover == Param::ADVANCE_BORIS) {
        // Boris push, MCC trial, boundary and deposit in one pass
        run_pass(M2D_STEP_PUSH | M2D_STEP_BOUNDARY | (p_->selfconsistent ? M2D_STEP_DEPOSIT : 0u));
    } else if (p_->mover == Param::ADVANCE_MULTICOLL) {
        ensure_device();
        if (cap_ > 0) {
            ensure_ttd();
            bind_partners();
            m2d_species2d d = desc();
            m2d_particles2d s = store();
            check(m2d_advance_multicoll_2d(dev_->ctx(), field->handle(), &d, &s, (double *)ttd_.get(), mcc_,
                                           nullptr));
            dev_newer_ = true;
        }
        advance_boundary();
        // advance_multicoll counts the step itself as well (particles.cpp:858-859)
        niter++;
        t += dt;
    } else {
        throw std::runtime_error("Species<CARTESIAN>: advance method not implemented");
    }
    niter++;
    t += dt;
}

void Species::advance_boundary()
{
    run_pass(M2D_STEP_BOUNDARY | (p_->selfconsistent ? M2D_STEP_DEPOSIT : 0u));
}

void Species::advance_init()
{
    // advance_position_init (particles.cpp:795-811): nothing for multicoll
    if (p_->mover == Param::ADVANCE_MULTICOLL) return;
    if (p_->mover != Param::ADVANCE_BORIS)
        throw std::runtime_error("Species<CARTESIAN>: advance method not implemented");
    ensure_device();
    if (cap_ == 0) return;
    m2d_species2d d = desc();
    m2d_particles2d s = store();
    check(m2d_advance_init_2d(dev_->ctx(), field->handle(), &d, &s));
    dev_newer_ = true;
}

void Species::accumulate()
{
    ensure_device();
    m2d_particles2d s = store();
    check(m2d_accumulate_2d(dev_->ctx(), &rho.grid(), &s, charge, rho.device()));
    rho.mark_device_written();
}

// Sample n particles into a fresh store with `fill`, then append them to
// this one.  `fill` gets the new store and, for the multicoll mover, the new
// slots' time_to_death (set to 0 before, as in a freshly inserted
// t_particle), else null.
void Species::append_samples(int nparticles, const std::function<void(const m2d_particles2d &, double *)> &fill)
{
    if (nparticles <= 0) return;
    ensure_device();
    DeviceBuffer tmp[5];
    for (auto &b : tmp) b = DeviceBuffer(dev_, (size_t)nparticles * sizeof(float));
    m2d_particles2d t{(float *)tmp[0].get(), (float *)tmp[1].get(), (float *)tmp[2].get(),
                      (float *)tmp[3].get(), (float *)tmp[4].get(), nparticles};
    const int64_t old = cap_, n = cap_ + nparticles;
    DeviceBuffer ttd;
    if (multicoll()) {
        ttd = DeviceBuffer(dev_, (size_t)n * sizeof(double));
        check(m2d_memset(dev_->ctx(), ttd.get(), 0, (size_t)n * sizeof(double)));
        if (old && ttd_.bytes() == (size_t)old * sizeof(double))
            check(m2d_memcpy_d2d(dev_->ctx(), ttd.get(), ttd_.get(), (size_t)old * sizeof(double)));
    }
    fill(t, multicoll() ? (double *)ttd.get() + old : nullptr);
    DeviceBuffer grown[5];
    for (int c = 0; c < 5; c++) {
        grown[c] = DeviceBuffer(dev_, (size_t)n * sizeof(float));
        if (old) check(m2d_memcpy_d2d(dev_->ctx(), grown[c].get(), soa_[c].get(), (size_t)old * sizeof(float)));
        check(m2d_memcpy_d2d(dev_->ctx(), (float *)grown[c].get() + old, tmp[c].get(),
                             (size_t)nparticles * sizeof(float)));
    }
    dev_->synchronize();
    for (int c = 0; c < 5; c++) soa_[c] = std::move(grown[c]);
    if (multicoll()) ttd_ = std::move(ttd);
    cap_ = n;
    dev_newer_ = true;
}

// each sampler call draws from its own stream
uint32_t Species::next_sample_stream() { return stream_ ^ (disk_calls_++ << 16); }

void Species::add_particles_on_disk(int nparticles, double centerx, double centery, double radius)
{
    const double vth = v_max * M_SQRT1_2;  // rndv: rnor() v_max / sqrt(2)
    append_samples(nparticles, [&](const m2d_particles2d &t, double *) {
        check(m2d_add_particles_on_disk_2d(dev_->ctx(), field->handle(), &t, centerx, centery, radius, vth,
                                           next_sample_stream()));
    });
}

void Species::add_particles_bessel(int nparticles, double centerx, double centery, double radius)
{
    const double vth = v_max * M_SQRT1_2;
    append_samples(nparticles, [&](const m2d_particles2d &t, double *) {
        check(m2d_add_particles_bessel_2d(dev_->ctx(), field->handle(), &t, centerx, centery, radius, vth,
                                          next_sample_stream()));
    });
}

void Species::add_particles_everywhere(int nparticles)
{
    const double vth = v_max * M_SQRT1_2;  // rnor() v_max / M_SQRT2 (particles.cpp:694-696)
    append_samples(nparticles, [&](const m2d_particles2d &t, double *ttd) {
        check(m2d_add_particles_everywhere_2d(dev_->ctx(), field->handle(), &t, vth, ttd, lifetime,
                                              next_sample_stream()));
    });
}

void Species::lifetime_init(const std::vector<Species *> &speclist)
{
    if (mcc_) {
        m2d_mcc_destroy(mcc_);
        mcc_ = nullptr;
    }
    partners_.clear();
    lifetime = INFINITY;
    if (!params_.interactions.empty()) build_mcc(speclist);
    // time_to_death of the live particles (particles.cpp:163-169); only the
    // multicoll mover reads it
    if (multicoll()) {
        ensure_device();
        if (cap_ > 0) {
            ensure_ttd();
            m2d_species2d d = desc();
            m2d_particles2d s = store();
            check(m2d_time_to_death_init_2d(dev_->ctx(), &d, &s, (double *)ttd_.get()));
            dev_newer_ = true;
        }
    }
}

void Species::build_mcc(const std::vector<Species *> &speclist)
{
    std::vector<m2d_target> targets;
    std::vector<m2d_interaction> inter;
    std::vector<double> tabE, tabS;
    std::vector<std::string> names;
    for (const InteractionParams &ip : params_.interactions) {
        auto it = std::find(names.begin(), names.end(), ip.target);
        int k = (int)(it - names.begin());
        if (it == names.end()) {
            Species *ts = nullptr;
            for (Species *s : speclist)
                if (s->name == ip.target) ts = s;
            if (!ts) throw std::runtime_error("Species " + name + ": unknown interaction target " + ip.target);
            m2d_target tg{};
            tg.mass = ts->mass;
            tg.density = ts->density;
            tg.temperature = ts->temperature;
            tg.charge = ts->charge;
            targets.push_back(tg);
            names.push_back(ip.target);
            partners_.push_back(ts->particle ? ts : nullptr);
        }
        m2d_interaction in{};
        in.type = ip.type;
        in.target = k;
        in.DE = ip.DE;
        in.rate = ip.rate;
        in.cutoff = ip.cutoff;
        if (!ip.cs_E.empty()) {
            if (ip.cs_E.size() != ip.cs_sigma.size())
                throw std::runtime_error("Species " + name + ": cross-section table columns differ");
            in.tab_offset = (int32_t)tabE.size();
            in.tab_len = (int32_t)ip.cs_E.size();
            tabE.insert(tabE.end(), ip.cs_E.begin(), ip.cs_E.end());
            tabS.insert(tabS.end(), ip.cs_sigma.begin(), ip.cs_sigma.end());
        }
        inter.push_back(in);
    }
    m2d_mcc_desc d{};
    d.mass = mass;
    d.charge = charge;
    d.temperature = temperature;
    d.density = density;
    d.E_max = E_max;
    d.n_targets = (int32_t)targets.size();
    d.targets = targets.data();
    d.n_inter = (int32_t)inter.size();
    d.inter = inter.data();
    d.n_tab = (int32_t)tabE.size();
    d.tab_E = tabE.empty() ? nullptr : tabE.data();
    d.tab_sigma = tabS.empty() ? nullptr : tabS.data();
    check(m2d_mcc_create(dev_->ctx(), &d, &mcc_));
    lifetime = m2d_mcc_lifetime(mcc_);
}

double Species::mean_energy()
{
    ensure_device();
    m2d_particles2d s = store();
    double mean = 0.0;
    int64_t live = 0;
    check(m2d_mean_energy_2d(dev_->ctx(), &s, mass, &mean, &live));
    return mean;
}

void Species::sort()
{
    if (multicoll())
        throw std::runtime_error("Species::sort: the multicoll mover keeps time_to_death per slot");
    ensure_device();
    if (cap_ == 0) return;
    for (int c = 0; c < 5; c++)
        if (tmp_[c].bytes() != (size_t)cap_ * sizeof(float))
            tmp_[c] = DeviceBuffer(dev_, (size_t)cap_ * sizeof(float));
    m2d_particles2d src = store();
    m2d_particles2d dst{(float *)tmp_[0].get(), (float *)tmp_[1].get(), (float *)tmp_[2].get(),
                        (float *)tmp_[3].get(), (float *)tmp_[4].get(), cap_};
    check(m2d_sort_2d(dev_->ctx(), field->handle(), &src, &dst, nullptr));
    for (int c = 0; c < 5; c++) std::swap(soa_[c], tmp_[c]);
    dev_newer_ = true;
}

// ================================================================ Pic
Pic::Pic(Param &p, const std::vector<SpeciesParams> &species, const t_grid *grid, int device)
    : param(p), dev(std::make_shared<Device>(device, p.seed)), field(p, dev, grid ? *grid : t_grid(p))
{
    p.check();
    for (size_t i = 0; i < species.size(); i++)
        speclist.emplace_back(new Species(species[i], p, field, (uint32_t)i));
    std::vector<Species *> all;
    for (auto &s : speclist) all.push_back(s.get());
    for (auto &s : speclist)
        if (s->particle) s->lifetime_init(all);
    // the constructor body of pic.cpp:130-188
    field.reset();
    if (p.particle_reload)
        for (auto &s : speclist)
            if (s->particle) s->load(p.particle_reload_dir + "/particles_" + s->name + ".dat");
    if (!p.magnetic_field_const) field.load_magnetic_field(p.magnetic_field_file.c_str());
    if (p.electric_field_from_file) {
        // the grid stays the Param's: Field2D::load checks that the files hold it
        field.u.load(p.electric_field_static_file.c_str());
        if (p.rf) field.uRF.load(p.electric_field_rf_file.c_str());
        field.update();
    } else {
        // the RF basis is solved for self-consistent runs too: it is not
        // recomputed afterwards
        field.boundary_solve_rf();
        if (!p.selfconsistent) {
            field.boundary_solve();
            field.reset();
        }
    }
}

Species *Pic::find(const std::string &name)
{
    for (auto &s : speclist)
        if (s->name == name) return s.get();
    return nullptr;
}

void Pic::run_initscript(const std::string &filename)
{
    std::ifstream fr(filename.c_str());
    if (!fr.is_open()) throw std::runtime_error("Pic::run_initscript: cannot open " + filename);
    std::string line;
    while (std::getline(fr, line)) {
        if (!line.empty() && line[0] == '#') continue;
        std::istringstream s(line);
        std::vector<std::string> toks;
        std::string tok;
        while (s >> tok) toks.push_back(tok);
        if (toks.empty()) continue;
        auto species = [&](size_t want) {
            if (toks.size() != want)
                throw std::runtime_error("Pic::run_initscript: wrong number of parameters (" +
                                         std::to_string(toks.size()) + ") to " + toks[0] + "\n");
            Species *sp = find(toks[1]);
            if (!sp) throw std::runtime_error("Pic::run_initscript: unrecognized species type \"" + toks[1] + "\"\n");
            return sp;
        };
        if (toks[0] == "add_particles_on_disk") {
            Species *sp = species(6);
            sp->add_particles_on_disk(std::stoi(toks[2]), std::stod(toks[3]), std::stod(toks[4]),
                                      std::stod(toks[5]));
        } else if (toks[0] == "add_tracked_particle") {
            // one particle at a given state (tracking output is not part of this build)
            Species *sp = species(7);
            int ii = sp->insert();
            t_particle &p = sp->particles[(size_t)ii];
            p.x = std::stod(toks[2]);
            p.z = std::stod(toks[3]);
            p.y = 0;
            p.vx = std::stod(toks[4]);
            p.vy = std::stod(toks[5]);
            p.vz = std::stod(toks[6]);
        } else if (toks[0] == "add_particles_bessel") {
            Species *sp = species(6);
            sp->add_particles_bessel(std::stoi(toks[2]), std::stod(toks[3]), std::stod(toks[4]),
                                     std::stod(toks[5]));
        } else if (toks[0] == "add_particles_everywhere") {
            Species *sp = species(3);
            sp->add_particles_everywhere(std::stoi(toks[2]));
        }
    }
}

void Pic::advance()
{
    if (param.selfconsistent) {
        field.boundary_solve();
        if (param.u_smooth) field.u_smooth();
        field.reset();
        for (auto &s : speclist)
            if (s->particle && s->n_particles() > 0) s->rho.reset();
    }
    for (auto &s : speclist) s->advance();
    for (auto &s : speclist)
        if (s->particle && param.selfconsistent) field.rho.add(s->rho);
    iter++;
}

void Pic::advance_init()
{
    if (param.selfconsistent) {
        for (auto &s : speclist)
            if (s->particle && s->n_particles() > 0) {
                s->accumulate();
                field.rho.add(s->rho);
            }
        field.boundary_solve();
        if (param.u_smooth) field.u_smooth();
        // boundary_solve builds its rhs in place on rho: rebuild rho
        field.reset();
        for (auto &s : speclist)
            if (s->particle && s->n_particles() > 0) field.rho.add(s->rho);
    }
    for (auto &s : speclist) s->advance_init();
}

}  // namespace mag2d

// ================================================================ C entry points
using namespace mag2d;

extern "C" {

// Host-only check of the slot container: insert/remove/resize semantics of
// BaseSpecies and the checkpoint round trip.  0 on success, else writes the
// failed check into msg.
int m2dh_list_selftest(const char *tmpfile, char *msg, int len)
{
    auto fail = [&](const char *what) {
        std::snprintf(msg, (size_t)len, "%s", what);
        return 1;
    };
    try {
        ParticleList L;
        if (L.n_particles() != 0) return fail("empty list has particles");
        int a = L.insert();  // grows to 0 * 1.1 + 1 = 1
        if (a != 0 || L.particles.size() != 1) return fail("first insert");
        int b = L.insert();  // grows to 1 * 1.1 + 1 = 2
        if (b != 1 || L.particles.size() != 2) return fail("second insert");
        for (int i = 0; i < 8; i++) L.insert();
        if (L.n_particles() != 10) return fail("count after 10 inserts");
        L.remove(3);
        bool threw = false;
        try {
            L.remove(3);
        } catch (const std::runtime_error &) {
            threw = true;
        }
        if (!threw) return fail("double remove must throw");
        if (L.insert() != 3) return fail("insert must reuse the last removed slot");
        size_t before = L.particles.size();
        L.resize((unsigned)before + 5);
        if (L.particles.size() != before + 5 || L.n_particles() != 10) return fail("resize");
        L.resize(2);
        if (L.particles.size() != before + 5) return fail("resize must not shrink");
        for (int i = 0; i < (int)L.particles.size(); i++)
            if (!L.particles[i].empty) {
                L.particles[i].x = i * 1e-3;
                L.particles[i].vz = -i;
            }
        L.remove(5);
        L.save(tmpfile);
        ParticleList M;
        M.load(tmpfile);
        if (M.particles.size() != L.particles.size() || M.n_particles() != L.n_particles())
            return fail("load: sizes");
        int k = 0;
        for (size_t i = 0; i < L.particles.size(); i++)
            if (!L.particles[i].empty) {
                if (M.particles[k].x != L.particles[i].x || M.particles[k].vz != L.particles[i].vz)
                    return fail("load: particle data");
                k++;
            }
        for (size_t i = (size_t)k; i < M.particles.size(); i++)
            if (!M.particles[i].empty) return fail("load: tail must be empty");
        if (M.insert() != (int)M.particles.size() - 1) return fail("load: free stack order");
        if (sizeof(t_particle) != 64) return fail("t_particle must be 64 bytes");
    } catch (const std::exception &e) {
        std::snprintf(msg, (size_t)len, "exception: %s", e.what());
        return 2;
    }
    msg[0] = 0;
    return 0;
}

// A trap run through Pic: EMPTY geometry (grounded frame), end caps along z
// at end_cap_V, uniform Bz, one species loaded on a disk, advance_init and
// `steps` advances.  Returns 0 or a negative status; positions and velocities
// of the n slots (live particles first, in store order) go to out_* when
// given, and metrics[0..5] = mean energy (eV), live count, sum of Fields::rho
// (C), u at the centre node (V), last solve's FP64 relative residual, its
// outer iterations.
typedef struct {
    int32_t grid, n_particles, steps, selfconsistent;
    double length, dt, Bz, end_cap_V, temperature_K, mass, charge, radius, density_total;
    uint64_t seed;
    int32_t u_smooth;  // Param::u_smooth
} m2dh_trap_cfg;

static int trap_run(const m2dh_trap_cfg *c, const char *bfile, float *out_x, float *out_z, float *out_vx,
                    float *out_vy, float *out_vz, int64_t n_out, double *metrics, char *msg, int len)
{
    try {
        Param p;
        p.x_sampl = p.z_sampl = c->grid;
        p.x_max = p.z_max = c->length;
        p.geometry = EMPTY;
        p.selfconsistent = c->selfconsistent;
        p.Bz = c->Bz;
        if (bfile) {
            p.magnetic_field_const = 0;
            p.magnetic_field_file = bfile;
        }
        p.n_particles_total = c->n_particles;
        p.density_total = c->density_total;
        p.seed = c->seed;
        p.u_smooth = c->u_smooth != 0;
        p.derive();
        SpeciesParams sp;
        sp.name = "particles";
        sp.mass = c->mass;
        sp.charge = c->charge;
        sp.temperature = c->temperature_K;
        sp.dt = c->dt;
        // grounded frame (EMPTY builder) with end caps on the first and last
        // two node rows along z
        t_grid g(p);
        const double h = 1.5 * p.dz;
        g.square_electrode(0, p.x_max, 0, h, c->end_cap_V);
        g.square_electrode(0, p.x_max, p.z_max - h, p.z_max, c->end_cap_V);
        Pic run(p, {sp}, &g);
        Species &s = *run.speclist[0];
        s.add_particles_on_disk(c->n_particles, 0.5 * p.x_max, 0.5 * p.z_max, c->radius);
        run.advance_init();
        for (int k = 0; k < c->steps; k++) run.advance();
        s.download();
        if (metrics) {
            metrics[0] = s.mean_energy();
            metrics[1] = s.n_particles();
            const Field2D &rho = run.field.rho, &u = run.field.u;
            double rs = 0;
            for (int i = 0; i < p.x_sampl; i++)
                for (int j = 0; j < p.z_sampl; j++) rs += rho[i][j];
            metrics[2] = rs;
            metrics[3] = u[p.x_sampl / 2][p.z_sampl / 2];
            metrics[4] = run.field.last_solve().rel_residual;
            metrics[5] = run.field.last_solve().outer_iters;
        }
        int64_t k = 0;
        for (const t_particle &q : s.particles) {
            if (q.empty || k >= n_out) continue;
            if (out_x) out_x[k] = (float)q.x;
            if (out_z) out_z[k] = (float)q.z;
            if (out_vx) out_vx[k] = (float)q.vx;
            if (out_vy) out_vy[k] = (float)q.vy;
            if (out_vz) out_vz[k] = (float)q.vz;
            k++;
        }
    } catch (const std::exception &e) {
        std::snprintf(msg, (size_t)len, "%s", e.what());
        return -1;
    }
    msg[0] = 0;
    return 0;
}

// Pic::run_initscript (pic.cpp:241-326) on a box of grid^2 nodes and side
// `length` (EMPTY geometry) with one species "ION" of the given mass and
// temperature.  Writes the live particles in slot order (up to n_out) and
// returns their count in *n_live.
int m2dh_initscript_run(const char *script, int32_t grid, double length, double mass, double temperature_K,
                        uint64_t seed, float *out_x, float *out_z, float *out_vx, float *out_vy, float *out_vz,
                        int64_t n_out, int64_t *n_live, char *msg, int len)
{
    try {
        Param p;
        p.x_sampl = p.z_sampl = grid;
        p.x_max = p.z_max = length;
        p.geometry = EMPTY;
        p.seed = seed;
        p.derive();
        SpeciesParams ion;
        ion.name = "ION";
        ion.mass = mass;
        ion.charge = physconst::q_e;
        ion.temperature = temperature_K;
        Pic run(p, {ion});
        run.run_initscript(script);
        Species &s = *run.speclist[0];
        s.download();
        int64_t k = 0;
        for (const t_particle &q : s.particles) {
            if (q.empty) continue;
            if (k < n_out) {
                out_x[k] = (float)q.x;
                out_z[k] = (float)q.z;
                out_vx[k] = (float)q.vx;
                out_vy[k] = (float)q.vy;
                out_vz[k] = (float)q.vz;
            }
            k++;
        }
        *n_live = k;
    } catch (const std::exception &e) {
        std::snprintf(msg, (size_t)len, "%s", e.what());
        return -1;
    }
    msg[0] = 0;
    return 0;
}

int m2dh_trap_run(const m2dh_trap_cfg *c, float *out_x, float *out_z, float *out_vx, float *out_vy,
                  float *out_vz, int64_t n_out, double *metrics, char *msg, int len)
{
    return trap_run(c, nullptr, out_x, out_z, out_vx, out_vy, out_vz, n_out, metrics, msg, len);
}

// the same trap with magnetic_field_const = 0 and B from `bfile` (the
// reference's "r z Br Bz" lines); cfg->Bz is not used
int m2dh_trap_run_bfile(const m2dh_trap_cfg *c, const char *bfile, float *out_x, float *out_z, float *out_vx,
                        float *out_vy, float *out_vz, int64_t n_out, double *metrics, char *msg, int len)
{
    return trap_run(c, bfile, out_x, out_z, out_vx, out_vy, out_vz, n_out, metrics, msg, len);
}

typedef struct {
    int32_t grid, n_particles, steps, mover;  // mover: Param::Mover
    double length, dt, extern_field, temperature_K, mass, charge, radius;
    double gas_mass, gas_temperature_K, gas_density, rate;  // one ELASTIC interaction with a background gas
    uint64_t seed;
} m2dh_mcc_cfg;

// A run in the style of config_MCC.txt: EMPTY geometry, constant
// extern_field, not self-consistent, one particle species that collides
// elastically with a background gas (a species that is not a particle
// ensemble), advanced by Pic::advance with the mover of cfg->mover.  Writes
// the live particles (slot order) and their time_to_death, and metrics =
// {mean energy eV, live particles, niter, lifetime}.
int m2dh_mcc_run(const m2dh_mcc_cfg *c, float *out_x, float *out_z, float *out_vx, float *out_vy, float *out_vz,
                 double *out_ttd, int64_t n_out, double *metrics, char *msg, int len)
{
    try {
        Param p;
        p.x_sampl = p.z_sampl = c->grid;
        p.x_max = p.z_max = c->length;
        p.geometry = EMPTY;
        p.selfconsistent = 0;
        p.extern_field = c->extern_field;
        p.mover = (Param::Mover)c->mover;
        p.n_particles_total = c->n_particles;
        p.seed = c->seed;
        p.derive();
        SpeciesParams ion;
        ion.name = "ION";
        ion.mass = c->mass;
        ion.charge = c->charge;
        ion.temperature = c->temperature_K;
        ion.dt = c->dt;
        InteractionParams ip;
        ip.type = M2D_ELASTIC;
        ip.target = "GAS";
        ip.rate = c->rate;
        ion.interactions.push_back(ip);
        SpeciesParams gas;
        gas.name = "GAS";
        gas.mass = c->gas_mass;
        gas.temperature = c->gas_temperature_K;
        gas.density = c->gas_density;
        gas.dt = c->dt;
        gas.particle = false;
        Pic run(p, {ion, gas});
        Species &s = *run.speclist[0];
        s.add_particles_on_disk(c->n_particles, 0.5 * p.x_max, 0.5 * p.z_max, c->radius);
        run.advance_init();
        for (int k = 0; k < c->steps; k++) run.advance();
        s.download();
        if (metrics) {
            metrics[0] = s.mean_energy();
            metrics[1] = s.n_particles();
            metrics[2] = (double)s.niter;
            metrics[3] = s.lifetime;
        }
        int64_t k = 0;
        for (const t_particle &q : s.particles) {
            if (q.empty || k >= n_out) continue;
            if (out_x) out_x[k] = (float)q.x;
            if (out_z) out_z[k] = (float)q.z;
            if (out_vx) out_vx[k] = (float)q.vx;
            if (out_vy) out_vy[k] = (float)q.vy;
            if (out_vz) out_vz[k] = (float)q.vz;
            if (out_ttd) out_ttd[k] = q.time_to_death;
            k++;
        }
    } catch (const std::exception &e) {
        std::snprintf(msg, (size_t)len, "%s", e.what());
        return -1;
    }
    msg[0] = 0;
    return 0;
}

// Field2D::print's text (Field2D.cpp:30-44) of host data (nx x nz,
// row-major) to a file
int m2dh_field_text_write(int32_t nx, int32_t nz, double dx, double dz, double xmin, double zmin,
                          const double *data, double factor, const char *path, char *msg, int len)
{
    try {
        std::ofstream out(path);
        if (!out.is_open()) throw std::runtime_error("Field2D::print(): failed opening file");
        field_text_write(out, m2d_grid2d{nx, nz, dx, dz, xmin, zmin}, data, factor);
        if (!out.good()) throw std::runtime_error("Field2D::print(): write error");
    } catch (const std::exception &e) {
        std::snprintf(msg, (size_t)len, "%s", e.what());
        return -1;
    }
    msg[0] = 0;
    return 0;
}

// Field2D::load's reader (Field2D.cpp:46-130): meta = {jmax, lmax, dx, dy,
// xmin, ymin}; the first min(cap, jmax lmax) values go to data.  -1 where the
// reference throws.
int m2dh_field_text_read(const char *path, double *meta, double *data, int64_t cap, char *msg, int len)
{
    try {
        FieldText t = field_text_read(path);
        meta[0] = t.g.nx;
        meta[1] = t.g.nz;
        meta[2] = t.g.dx;
        meta[3] = t.g.dz;
        meta[4] = t.g.xmin;
        meta[5] = t.g.zmin;
        const size_t n = std::min<size_t>(t.data.size(), cap > 0 ? (size_t)cap : 0);
        if (n) std::memcpy(data, t.data.data(), n * sizeof(double));
    } catch (const std::exception &e) {
        std::snprintf(msg, (size_t)len, "%s", e.what());
        return -1;
    }
    msg[0] = 0;
    return 0;
}

// BaseSpecies::load (particles.cpp:57-93) on the host: the checkpoint's
// slots (live ones first) as 64-byte t_particle records into out (up to cap
// slots), after the re-draw of positions outside the box
int m2dh_particles_load(const char *path, double x_min, double x_max, double z_min, double z_max, uint64_t seed,
                        uint32_t stream, void *out, int64_t cap, int64_t *n_slots, int64_t *n_live, char *msg,
                        int len)
{
    try {
        ParticleList L;
        L.load(path);
        const int live = L.n_particles();
        reload_remap(L.particles, live, x_min, x_max, z_min, z_max, seed, stream);
        *n_slots = (int64_t)L.particles.size();
        *n_live = live;
        const size_t n = std::min<size_t>(L.particles.size(), cap > 0 ? (size_t)cap : 0);
        if (n) std::memcpy(out, L.particles.data(), n * sizeof(t_particle));
    } catch (const std::exception &e) {
        std::snprintf(msg, (size_t)len, "%s", e.what());
        return -1;
    }
    msg[0] = 0;
    return 0;
}

// Param::check (param.cpp:60-64) of the given switches: 0, or -1 and the
// reference's message
int m2dh_param_check(int32_t selfconsistent, int32_t rf, int32_t electric_field_from_file, char *msg, int len)
{
    try {
        Param p;
        p.selfconsistent = selfconsistent;
        p.rf = rf;
        p.electric_field_from_file = electric_field_from_file;
        p.check();
    } catch (const std::exception &e) {
        std::snprintf(msg, (size_t)len, "%s", e.what());
        return -1;
    }
    msg[0] = 0;
    return 0;
}

typedef struct {
    int32_t grid, steps, geometry, rf;
    double length, dt, rf_amplitude, rf_omega, mass, charge;
    uint64_t seed;
} m2dh_file_cfg;

// A run started from files, as Pic's constructor reads them
// (pic.cpp:138-175): one particle species "ION" reloaded from
// reload_dir/particles_ION.dat, u (and uRF when cfg->rf) loaded from text
// files in Field2D::print's format, no solve (electric_field_from_file,
// not self-consistent); then Pic::advance_init and cfg->steps Pic::advance.
// Writes the live particles in slot order (up to n_out) and their count.
int m2dh_file_run(const m2dh_file_cfg *c, const char *u_file, const char *urf_file, const char *reload_dir,
                  float *out_x, float *out_z, float *out_vx, float *out_vy, float *out_vz, int64_t n_out,
                  int64_t *n_live, char *msg, int len)
{
    try {
        Param p;
        p.x_sampl = p.z_sampl = c->grid;
        p.x_max = p.z_max = c->length;
        p.geometry = (geometry_type)c->geometry;
        p.selfconsistent = 0;
        p.rf = c->rf;
        p.rf_amplitude = c->rf_amplitude;
        p.rf_omega = c->rf_omega;
        p.electric_field_from_file = 1;
        p.electric_field_static_file = u_file;
        if (urf_file) p.electric_field_rf_file = urf_file;
        p.particle_reload = 1;
        p.particle_reload_dir = reload_dir;
        p.seed = c->seed;
        p.derive();
        SpeciesParams ion;
        ion.name = "ION";
        ion.mass = c->mass;
        ion.charge = c->charge;
        ion.dt = c->dt;
        Pic run(p, {ion});
        Species &s = *run.speclist[0];
        run.advance_init();
        for (int k = 0; k < c->steps; k++) run.advance();
        s.download();
        int64_t k = 0;
        for (const t_particle &q : s.particles) {
            if (q.empty) continue;
            if (k < n_out) {
                out_x[k] = (float)q.x;
                out_z[k] = (float)q.z;
                out_vx[k] = (float)q.vx;
                out_vy[k] = (float)q.vy;
                out_vz[k] = (float)q.vz;
            }
            k++;
        }
        *n_live = k;
    } catch (const std::exception &e) {
        std::snprintf(msg, (size_t)len, "%s", e.what());
        return -1;
    }
    msg[0] = 0;
    return 0;
}

}  // extern "C"
