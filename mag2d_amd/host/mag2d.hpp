// mag2d.hpp -- the reference's C++ class API (rouckas/mag2d) rebuilt over the
// C ABI of libmag2d_amd (include/mag2d_amd.h).
//
// Names, argument meaning and error behaviour follow the reference:
//   t_particle, BaseSpecies container        particles.hpp:26-33, 100-131, 223-249
//   Species<CARTESIAN>::advance & co.        particles.hpp:281-291, 342-411
//   Field2D                                  Field2D.hpp:14-168, Field2D.cpp:30-130
//   t_grid, Fields                           fields.hpp:39-177, fields.cpp:115-360
//   Param (the members the loop reads)       param.hpp, param.cpp:14-142
//   Pic<CARTESIAN>::advance / advance_init   pic.cpp:330-384
// Every failure of the library is rethrown as std::runtime_error, as the
// reference throws.  Compute runs on the GPU only: there is no host path.
//
// Host mirrors.  Field2D and the particle vector keep a host copy next to the
// device data.  operator[] and download() bring the device data to the host;
// writes through operator[] / insert() / remove() / resize() are sent back
// before the next device call.  Writing `particles[j]` directly must be
// followed by upload(), as the vector itself cannot see the write.
#pragma once

#include <cstdint>
#include <functional>
#include <iosfwd>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "mag2d_amd.h"

namespace mag2d {

// physical constants of the reference (param.cpp:7-11)
namespace physconst {
constexpr double k_B = 1.380662e-23;
constexpr double q_e = 1.602189e-19;
constexpr double epsilon_0 = 8.854187817e-12;
}  // namespace physconst

// status != M2D_OK -> std::runtime_error(m2d_last_error())
void check(int status);

// ------------------------------------------------------------------ device
// One GPU context (stream, scratch, counters).  Shared by the objects that
// use it; it outlives them.
class Device {
public:
    explicit Device(int device = 0, uint64_t seed = 1234);
    ~Device();
    Device(const Device &) = delete;
    Device &operator=(const Device &) = delete;
    m2d_ctx *ctx() const { return ctx_; }
    void synchronize();
    // device status flags (M2D_FLAG_*) raised since the last clear
    uint32_t status(bool clear = true);
    // small staging buffer for point queries (doubles)
    double *staging(size_t n_doubles);

private:
    m2d_ctx *ctx_ = nullptr;
    void *staging_ = nullptr;
    size_t staging_n_ = 0;
};

// device allocation owned by a host object
class DeviceBuffer {
public:
    DeviceBuffer() = default;
    DeviceBuffer(std::shared_ptr<Device> dev, size_t bytes);
    ~DeviceBuffer();
    DeviceBuffer(DeviceBuffer &&o) noexcept;
    DeviceBuffer &operator=(DeviceBuffer &&o) noexcept;
    DeviceBuffer(const DeviceBuffer &) = delete;
    DeviceBuffer &operator=(const DeviceBuffer &) = delete;
    void *get() const { return p_; }
    size_t bytes() const { return bytes_; }

private:
    std::shared_ptr<Device> dev_;
    void *p_ = nullptr;
    size_t bytes_ = 0;
};

// ------------------------------------------------------------------ particles
// t_particle (particles.hpp:26-33): 64 bytes, y and time_to_death carried
// through unchanged by the Cartesian loop
class t_particle {
public:
    double x, y, z;
    double vx, vy, vz;
    double time_to_death;
    bool empty;
};

// The slot container of BaseSpecies (particles.hpp:100-131, 223-249):
// particles with holes plus the LIFO stack of empty slots.  Host only.
class ParticleList {
public:
    std::vector<t_particle> particles;
    // particles.hpp:223-233: throws when the slot is already empty
    void remove(int n);
    // particles.hpp:235-246: grows by 10 % + 1 when no slot is free
    int insert();
    int n_particles() const { return (int)(particles.size() - empty_.size()); }
    // particles.cpp:95-107: only grows
    void resize(unsigned int newsize);
    // rebuild the free stack after the slots were rewritten wholesale
    void rebuild_free_list();
    // checkpoint (particles.cpp:31-93): int size, int n_live, n_live x t_particle
    void save(const std::string &filename) const;
    // load keeps the reference's layout: live particles first, then empty
    // slots.  Species::load adds the re-draw of positions outside the box.
    void load(const std::string &filename);

protected:
    std::vector<int> empty_;
};

// ------------------------------------------------------------------ Field2D
// Field2D (Field2D.hpp:14-42) with its data on the device.  jmax = x samples,
// lmax = z samples, data[i][j] row-major like Array2D.
class Field2D {
public:
    Field2D(std::shared_ptr<Device> dev, int x_sampl, int y_sampl, double dx, double dy,
            double xmin = 0, double ymin = 0);
    // a view of device storage owned elsewhere (Fields::u, uRF, rho)
    Field2D(std::shared_ptr<Device> dev, double *device_data, int x_sampl, int y_sampl, double dx,
            double dy, double xmin, double ymin);
    Field2D(const Field2D &) = delete;
    Field2D &operator=(const Field2D &) = delete;

    int jmax, lmax;
    double GetDx() const { return g_.dx; }
    double GetDy() const { return g_.dz; }
    double GetXMin() const { return g_.xmin; }
    double GetYMin() const { return g_.zmin; }
    const m2d_grid2d &grid() const { return g_; }

    // host rows (Array2D::operator[], Array.hpp:30-33); the non-const form
    // marks the host copy as written
    double *operator[](int i);
    const double *operator[](int i) const;
    void reset();                       // Array2D::reset
    void add(const Field2D &other);     // Array2D::add
    // Field2D::accumulate (Field2D.hpp:45-62); positions go through FP32 like
    // the particle store; throws "outside of range" like the reference
    void accumulate(double q, double x, double y);
    void accumulate(double q, const double *x, const double *y, size_t n);
    double interpolate(double x, double y);                          // Field2D.hpp:64-78
    void grad(double x, double y, double &gx, double &gy);           // Field2D.hpp:80-168
    void interpolate(const double *x, const double *y, size_t n, double *out);
    void grad(const double *x, const double *y, size_t n, double *gx, double *gy);
    // text format of Field2D::print / load (Field2D.cpp:30-130): "x z value" per
    // node, a blank line after each x row
    void print(std::ostream &out, double factor = 1.0);
    void print(const char *filename, double factor = 1.0);
    // the file must hold this field's grid: its shape, and its spacing and
    // origin to the 6 significant digits print writes
    void load(const char *filename);

    // explicit mirror control
    double *device();                   // device data, host writes sent first
    void download();
    void upload();
    void mark_device_written() { state_ = DEVICE_NEWER; }
    bool host_newer() const { return state_ == HOST_NEWER; }

private:
    enum State { SYNCED, HOST_NEWER, DEVICE_NEWER };
    std::shared_ptr<Device> dev_;
    m2d_grid2d g_;
    DeviceBuffer own_;
    double *d_ = nullptr;
    mutable std::vector<double> h_;
    mutable State state_ = SYNCED;
    void pull() const;
};

// The text format of Field2D::print / load (Field2D.cpp:30-130) on host data
// (field_text_read: Field2D::load's reader, any order of nodes)
struct FieldText {
    m2d_grid2d g;                 // jmax, lmax, dx, dy, xmin, ymin as found in the file
    std::vector<double> data;     // data[i * lmax + j]
};
void field_text_write(std::ostream &out, const m2d_grid2d &g, const double *data, double factor = 1.0);
FieldText field_text_read(const char *filename);

// the re-draw of BaseSpecies::load for slots [0, live) of a loaded list
// (particles.cpp:77-83), with the Philox stream described at Species::load
void reload_remap(std::vector<t_particle> &particles, int live, double x_min, double x_max, double z_min,
                  double z_max, uint64_t seed, uint32_t stream);

// ------------------------------------------------------------------ Param
enum geometry_type { EMPTY = M2D_GEO_EMPTY, PROBE = M2D_GEO_PROBE, RF_22PT = M2D_GEO_RF_22PT,
                     RF_8PT = M2D_GEO_RF_8PT, RF_HAITRAP = M2D_GEO_RF_HAITRAP, RF_QUAD = M2D_GEO_RF_QUAD,
                     MAC = M2D_GEO_MAC, PENNING = M2D_GEO_PENNING, PENNING_SIMPLE = M2D_GEO_PENNING_SIMPLE,
                     TUBE = M2D_GEO_TUBE };
enum mask_type { FIXED = M2D_FIXED, FIXED_RF = M2D_FIXED_RF, FREE = M2D_FREE, BOUNDARY = M2D_BOUNDARY };
enum coord_type { CARTESIAN = 0 };

// The Param members read by the loop, with the reference's defaults
// (param.cpp:14-142).  Call derive() after changing sizes or counts.
struct Param {
    enum boundary_type { FREE = M2D_BC_FREE, PERIODIC = M2D_BC_PERIODIC };
    // Param::Mover (param.hpp:21); the config accepts ADVANCE_BORIS, the
    // default, and ADVANCE_MULTICOLL (param.cpp:97-105)
    enum Mover { ADVANCE_BORIS, ADVANCE_MULTICOLL, ADVANCE_LEAPFROG };
    Mover mover = ADVANCE_BORIS;
    int x_sampl = 101, z_sampl = 101;
    double x_max = 1e-2, z_max = 1e-2, x_min = 0.0, z_min = 0.0;
    geometry_type geometry = EMPTY;
    double probe_radius = 1e-3, u_probe = 0.0, extern_field = 0.0;
    int selfconsistent = 0;
    // electric_field_from_file: Pic loads u (and uRF when rf) from text files
    // in the format of Field2D::print instead of solving them (param.cpp:35-37,
    // pic.cpp:154-175); the files must hold this grid
    int electric_field_from_file = 0;
    std::string electric_field_static_file, electric_field_rf_file;
    // particle_reload: Pic loads every particle species from
    // particle_reload_dir/particles_<name>.dat (param.cpp:73-74, pic.cpp:138-146)
    int particle_reload = 0;
    std::string particle_reload_dir = ".";
    int boundary = FREE;
    int rf = 0;
    double rf_amplitude = 0.0, rf_omega = 0.0, rf_U0 = 0.0;
    double Br = 0.0, Bz = 0.0, Bt = 0.0;
    // 0: B from magnetic_field_file, loaded by Pic (param.cpp:41-42, pic.cpp:148-149)
    int magnetic_field_const = 1;
    std::string magnetic_field_file;
    double n_particles_total = 1e5, density_total = 1e11;
    uint64_t seed = 1234;
    double solve_tolerance = 1e-10;  // FP64 relative residual of every solve
    bool u_smooth = false;           // Fields::u_smooth after every solve (param.cpp:53, pic.cpp:336,372)
    // derived (param.cpp:126-142)
    double dx = 0, dz = 0, idx = 0, idz = 0, V = 0, dV = 0;
    void derive();
    // param.cpp:60-64: selfconsistent excludes rf and electric_field_from_file
    void check() const;
};

// ------------------------------------------------------------------ t_grid
// t_grid (fields.hpp:39-55): node masks and electrode voltages, host arrays
class t_grid {
public:
    explicit t_grid(const Param &p);   // the builder of p.geometry (fields.cpp:374-833)
    int jmax, lmax;
    double dx, dz;
    std::vector<uint8_t> mask;          // mask[i * lmax + j]
    std::vector<double> voltage;
    void square_electrode(double rmin, double rmax, double zmin, double zmax, double v,
                          mask_type type = FIXED);                     // fields.cpp:835-850
    void circle_electrode(double x, double z, double r, double v,
                          mask_type type = FIXED);                     // fields.cpp:852-868
    bool is_free(double x, double z) const;                            // fields.hpp:94-101
    m2d_grid2d grid() const;
};

// ------------------------------------------------------------------ Fields
// Fields (fields.hpp:57-92): u, uRF, rho on the device and the Poisson solve
class Fields {
    // declared first: the field views below are built from the library object
    Param *p_;
    std::shared_ptr<Device> dev_;
    m2d_fields2d *f_ = nullptr;
    m2d_solve_info info_{};

public:
    Fields(Param &p, std::shared_ptr<Device> dev);
    // a geometry edited after its builder ran (square_electrode & co.)
    Fields(Param &p, std::shared_ptr<Device> dev, const t_grid &grid);
    ~Fields();
    Fields(const Fields &) = delete;
    Fields &operator=(const Fields &) = delete;

    t_grid grid;
    Field2D u, uRF, rho;
    // fields.cpp:278-312 (rho -> u), 314-348 (voltage basis -> uRF), 350-353, 355-360
    void boundary_solve();
    void boundary_solve_rf();
    void solve();
    void reset();
    // Fields::E / B at one point (fields.hpp:124-177), t in s
    void E(double x, double y, double &gx, double &gy, double t);
    void B(double x, double y, double &Br, double &Bz, double &Bt) const;
    // fields.cpp:870-959: lines "r z Br Bz" of a regular grid in any order of
    // rows; the passes then take B from it (magnetic_field_const = 0)
    void load_magnetic_field(const char *fname);
    // send host edits of u / uRF and rebuild the gather tables
    void update();
    // fields.cpp:28-113 on u (field == nullptr) or uRF
    void u_smooth(bool symmetry = false, double radius = -1., Field2D *field = nullptr);
    m2d_fields2d *handle();
    const m2d_solve_info &last_solve() const { return info_; }
    Param &param() { return *p_; }
    std::shared_ptr<Device> device() const { return dev_; }

private:
    void run_solve(int mode);
};

// ------------------------------------------------------------------ Species
struct InteractionParams {         // one INTERACTION block (parser.cpp:97-143)
    int type = M2D_ELASTIC;        // M2D_ELASTIC / LANGEVIN / CX / SUPERELASTIC / COULOMB (which
                                   // needs the species' temperature and density and both charges)
    std::string target;            // species name
    double rate = 0.0;             // RATE (raw; LANGEVIN scales by cutoff^2)
    double DE = 0.0;               // eV
    double cutoff = 0.0;
    std::vector<double> cs_E, cs_sigma;  // CROSS_SECTION table (eV, m^2), optional
};

struct SpeciesParams {             // SpeciesParams (parser.hpp)
    std::string name;
    double mass = 0.0, charge = 0.0, temperature = 300.0, density = 0.0, dt = 1e-9;
    double E_max = 0.0;            // <= 0: 10 k T / q_e (particles.hpp:160)
    bool particle = true;          // false: a neutral background continuum
    std::vector<InteractionParams> interactions;
};

// Species<CARTESIAN> (particles.hpp:281-349) over the fused particle pass
class Species : public ParticleList {
public:
    Species(const SpeciesParams &sp, Param &param, Fields &field, uint32_t stream);
    ~Species();
    Species(const Species &) = delete;
    Species &operator=(const Species &) = delete;

    std::string name;
    bool particle;
    double mass, charge, temperature, density, dt, t, lifetime, E_max, v_max;
    uint64_t niter = 0;
    Field2D rho;
    Fields *field;

    // the operator API (particles.hpp:281-291)
    void advance();           // advance_position + advance_boundary, niter++, t += dt
    void advance_init();      // advance_position_init: advance_boris_init, nothing for multicoll
    void accumulate();        // rho.accumulate over live particles
    void advance_boundary();  // box, is_free, deposit when selfconsistent
    // particles.cpp:728-749; velocities Maxwellian at the species temperature
    void add_particles_on_disk(int nparticles, double centerx, double centery, double radius);
    // particles.cpp:701-726: density J0(2.405 r / radius) on the disk
    void add_particles_bessel(int nparticles, double centerx, double centery, double radius);
    // particles.cpp:686-699: uniform in the box, time_to_death = rexp() lifetime
    void add_particles_everywhere(int nparticles);
    // BaseSpecies::lifetime_init (particles.cpp:142-170) over the species list
    void lifetime_init(const std::vector<Species *> &speclist);
    // energy_dist.mean_tot() (histogram.cpp:68-71): mean kinetic energy in eV
    double mean_energy();
    // re-order the device store by tile (drops holes; host mirror follows).
    // Boris only: the multicoll mover keeps time_to_death per slot.
    void sort();
    // BaseSpecies::load (particles.cpp:57-93): the checkpoint of save(); a
    // live particle whose x (z) lies outside [x_min, x_max] ([z_min, z_max])
    // is put at a uniform position along that axis.  The uniforms are the
    // first (x) and second (z) word of the Philox stream (seed, species
    // stream, slot, step 0, purpose 5), so a reload is reproducible.
    void load(const std::string &filename);
    // the container calls of BaseSpecies on the host mirror
    int insert();
    void remove(int n);
    void resize(unsigned int newsize);
    int n_particles();

    // mirror control and the raw store
    void download();
    void upload();
    m2d_particles2d store();
    const m2d_step_stats &last_stats() const { return stats_; }

private:
    Param *p_;
    std::shared_ptr<Device> dev_;
    uint32_t stream_;
    SpeciesParams params_;
    uint32_t disk_calls_ = 0;
    DeviceBuffer soa_[5], tmp_[5];
    DeviceBuffer ttd_;                  // time_to_death per slot (FP64), multicoll mover only
    int64_t cap_ = 0;
    bool host_newer_ = false, dev_newer_ = false;
    m2d_mcc *mcc_ = nullptr;
    std::vector<Species *> partners_;   // target index -> live species (or null)
    m2d_step_stats stats_{};
    void ensure_device();
    void ensure_host();
    void reserve_device(int64_t n);
    m2d_species2d desc() const;
    void run_pass(uint32_t flags);
    bool multicoll() const { return p_->mover == Param::ADVANCE_MULTICOLL; }
    void ensure_ttd();
    void bind_partners();
    void build_mcc(const std::vector<Species *> &speclist);
    void append_samples(int nparticles, const std::function<void(const m2d_particles2d &, double *)> &fill);
    uint32_t next_sample_stream();
};

// ------------------------------------------------------------------ Pic
// Pic<CARTESIAN> (pic.cpp:20-384): the step ordering of one PIC iteration
class Pic {
public:
    // grid: a geometry edited after its builder ran; nullptr builds t_grid(param)
    Pic(Param &param, const std::vector<SpeciesParams> &species, const t_grid *grid = nullptr,
        int device = 0);
    Param &param;
    std::shared_ptr<Device> dev;
    Fields field;
    std::vector<std::unique_ptr<Species>> speclist;
    unsigned long iter = 0;
    void check_params() const { param.check(); }
    // initscript commands (pic.cpp:241-326): add_particles_on_disk
    void run_initscript(const std::string &filename);
    void advance();        // pic.cpp:330-358
    void advance_init();   // pic.cpp:359-384
    Species *find(const std::string &name);
};

}  // namespace mag2d
