/*
 * inovesa_amd.hpp -- C++ interface in the shape of Inovesa's phase-space
 * classes, over the C ABI of vfps_gpu.h (header only, C++14).
 *
 * Class names, constructor arguments, method names and error behaviour
 * follow the reference:
 *   PhaseSpace        inc/PS/PhaseSpace.hpp:40-552
 *   SourceMap         inc/SM/SourceMap.hpp:53-87
 *   KickMap           inc/SM/KickMap.hpp:22-80 (swapOffset, getForce)
 *   DriftMap          inc/SM/DriftMap.hpp:17-25
 *   RFKickMap         inc/SM/RFKickMap.hpp:17-34
 *   RotationMap       inc/SM/RotationMap.hpp:40-45
 *   FokkerPlanckMap   inc/SM/FokkerPlanckMap.hpp:25-64
 *   Identity          inc/SM/Identity.hpp:14-61
 *   WakePotentialMap  inc/SM/WakePotentialMap.hpp:20-32
 *   ElectricField     inc/PS/ElectricField.hpp:49-103, 107-170
 *   Impedance & co.   inc/Z/ (all), makeImpedance inc/Z/ImpedanceFactory.hpp:35
 * The device handle `oclhptr_t` is a shared Device (a null handle selects
 * the process-wide default device: there is no CPU path).  Constructors
 * throw std::invalid_argument where the reference does and
 * inovesa_amd::Error for device failures.  Data lives on the device;
 * getData(), getProjection(), wakePotential() and the other accessors that
 * return host pointers copy from the device first and point into a host
 * mirror owned by the object, valid until the next such call.
 * syncCLMem(cpu2dev) uploads the host mirror after it was written through
 * getData().
 */
#ifndef INOVESA_AMD_HPP
#define INOVESA_AMD_HPP

#include <array>
#include <complex>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "vfps_gpu.h"

namespace inovesa_amd {

typedef uint32_t meshindex_t;
typedef float meshaxis_t;
typedef float meshdata_t;
typedef float frequency_t;
typedef float timeaxis_t;
typedef float csrpower_t;
typedef double integral_t;  // FP64 charge integral (the reference's is float)
typedef std::complex<float> impedance_t;

// ---- errors -------------------------------------------------------------
class Error : public std::runtime_error {
public:
    Error(int code, const std::string& msg)
        : std::runtime_error("vfps error " + std::to_string(code) + ": " + msg),
          code_(code) {}
    int code() const { return code_; }

private:
    int code_;
};

inline void check(int rc)
{
    if (rc == VFPS_OK) {
        return;
    }
    const char* m = vfps_last_error();
    const std::string msg = m ? m : "";
    if (rc == VFPS_ERR_INVALID) {
        throw std::invalid_argument(msg);
    }
    throw Error(rc, msg);
}

// ---- device (stands for OCLH, inc/CL/OpenCLHandler.hpp:63-314) ------------
class Device {
public:
    explicit Device(int ordinal = 0) { check(vfps_dev_open(ordinal, &h_)); }
    ~Device() { vfps_dev_close(h_); }
    Device(const Device&) = delete;
    Device& operator=(const Device&) = delete;
    vfps_dev* handle() const { return h_; }
    // OCLH::finish (src/main.cpp:1046-1050)
    void finish() { check(vfps_sync(h_)); }

    static std::shared_ptr<Device> defaultDevice()
    {
        static std::shared_ptr<Device> d = std::make_shared<Device>(0);
        return d;
    }

private:
    vfps_dev* h_ = nullptr;
};

typedef std::shared_ptr<Device> oclhptr_t;

inline oclhptr_t resolve(oclhptr_t oclh)
{
    return oclh ? oclh : Device::defaultDevice();
}

// ---- PhaseSpace -----------------------------------------------------------
class PhaseSpace {
public:
    enum class clCopyDirection { cpu2dev, dev2cpu };

    // mesh size shared by all phase spaces (PhaseSpace::setSize,
    // inc/PS/PhaseSpace.hpp:361-387)
    static void setSize(meshindex_t x, meshindex_t b) { sizes() = {x, b}; }
    static void resetSize(meshindex_t x, meshindex_t b) { setSize(x, b); }
    static meshindex_t nx() { return sizes()[0]; }
    static meshindex_t nb() { return sizes()[1]; }

    // PhaseSpace(qmin, qmax, qscale, pmin, pmax, pscale, oclh, Qb, Ib,
    // filling, zoom, data) (src/PS/PhaseSpace.cpp:149-168)
    PhaseSpace(meshaxis_t qmin, meshaxis_t qmax, double qscale,
               meshaxis_t pmin, meshaxis_t pmax, double pscale,
               oclhptr_t oclh, double beam_charge, double beam_current,
               const std::vector<float>& filling = {1}, double zoom = 1,
               const meshdata_t* data = nullptr)
        : dev_(resolve(oclh)), scale_{{qscale, pscale}}
    {
        check(vfps_mesh_create(dev_->handle(), (uint32_t)filling.size(), nx(),
                               qmin, qmax, qscale, pmin, pmax, pscale,
                               beam_charge, beam_current, filling.data(),
                               zoom, data, &h_));
        refresh_info();
    }

    // copy constructor (src/PS/PhaseSpace.cpp:170-180)
    PhaseSpace(const PhaseSpace& other) : dev_(other.dev_), scale_(other.scale_)
    {
        check(vfps_mesh_clone(other.h_, &h_));
        refresh_info();
    }

    ~PhaseSpace() { vfps_mesh_destroy(h_); }

    // operator= (src/PS/PhaseSpace.cpp:374-379): copies the mesh
    PhaseSpace& operator=(const PhaseSpace& other)
    {
        if (this != &other) {
            check(vfps_mesh_copy(h_, other.h_));
        }
        return *this;
    }

    vfps_mesh* handle() const { return h_; }
    const oclhptr_t& device() const { return dev_; }

    // getData (inc/PS/PhaseSpace.hpp:105-109): the mesh, copied from the
    // device into the host mirror
    meshdata_t* getData()
    {
        syncCLMem(clCopyDirection::dev2cpu);
        return host_.data();
    }

    // syncCLMem (src/PS/PhaseSpace.cpp:381-414)
    void syncCLMem(clCopyDirection dir)
    {
        host_.resize((size_t)nb_ * n_ * n_);
        if (dir == clCopyDirection::dev2cpu) {
            check(vfps_mesh_download(h_, host_.data()));
        } else {
            check(vfps_mesh_upload(h_, host_.data()));
        }
    }

    void updateXProjection() { check(vfps_mesh_project_x(h_)); }
    void updateYProjection() { check(vfps_mesh_project_y(h_)); }

    // getProjection(axis) ([nb][n])
    const std::vector<float>& getProjection(uint_fast8_t axis)
    {
        proj_[axis].resize((size_t)nb_ * n_);
        check(vfps_mesh_get_projection(h_, axis, proj_[axis].data()));
        return proj_[axis];
    }

    void setProjection(uint_fast8_t axis, const std::vector<float>& p)
    {
        if (p.size() != (size_t)nb_ * n_) {
            throw std::invalid_argument("setProjection: size mismatch");
        }
        check(vfps_mesh_set_projection(h_, axis, p.data()));
    }

    // integrate (src/PS/PhaseSpace.cpp:200-239), FP64
    void integrate()
    {
        fill_.resize(nb_);
        check(vfps_mesh_integrate(h_, fill_.data(), &integral_));
    }
    std::vector<integral_t> getBunchPopulation() const { return fill_; }
    integral_t getIntegral() const { return integral_; }
    // normalize (src/PS/PhaseSpace.cpp:342-372)
    void normalize() { check(vfps_mesh_normalize(h_)); }
    const std::vector<integral_t>& integrateAndNormalize()
    {
        integrate();
        normalize();
        return fill_;
    }

    // variance(axis) (src/PS/PhaseSpace.cpp:241-289): per bunch mean, var, rms
    void variance(uint_fast8_t axis)
    {
        mean_[axis].resize(nb_);
        var_[axis].resize(nb_);
        rms_[axis].resize(nb_);
        check(vfps_mesh_variance(h_, axis, mean_[axis].data(),
                                 var_[axis].data(), rms_[axis].data()));
    }
    float getMoment(uint_fast8_t axis, uint_fast8_t m, meshindex_t b = 0) const
    {
        return m == 0 ? mean_[axis].at(b) : var_[axis].at(b);
    }
    float getBunchLength(meshindex_t b = 0) const { return rms_[0].at(b); }
    float getEnergySpread(meshindex_t b = 0) const { return rms_[1].at(b); }

    meshaxis_t getDelta(uint_fast8_t axis) const { return delta_[axis]; }
    // getScale(axis): metres (axis 0) or electron volts (axis 1) per unit
    double getScale(uint_fast8_t axis) const { return scale_[axis]; }
    meshaxis_t getZerobin(uint_fast8_t axis) const { return zero_[axis]; }
    std::vector<float> getAxis(uint_fast8_t axis) const
    {
        std::vector<float> a(n_);
        check(vfps_mesh_axis(h_, axis, a.data()));
        return a;
    }
    std::vector<float> simpsonWeights() const
    {
        std::vector<float> w(n_);
        check(vfps_mesh_simpson_weights(h_, w.data()));
        return w;
    }
    meshindex_t nBunches() const { return nb_; }
    meshindex_t nMeshCells() const { return n_ * n_; }
    meshindex_t nMeshCells(uint_fast8_t) const { return n_; }

private:
    static std::array<meshindex_t, 2>& sizes()
    {
        static std::array<meshindex_t, 2> s = {{0, 1}};
        return s;
    }

    void refresh_info()
    {
        check(vfps_mesh_info(h_, &nb_, &n_, &delta_[0], &delta_[1], &zero_[0],
                             &zero_[1]));
    }

    oclhptr_t dev_;
    std::array<double, 2> scale_;
    vfps_mesh* h_ = nullptr;
    uint32_t nb_ = 1, n_ = 0;
    float delta_[2] = {0, 0}, zero_[2] = {0, 0};
    std::vector<meshdata_t> host_;
    std::vector<float> proj_[2];
    std::vector<integral_t> fill_;
    integral_t integral_ = 0;
    std::vector<float> mean_[2], var_[2], rms_[2];
};

// ---- impedances (inc/Z/*.hpp) -------------------------------------------------
class Impedance {
public:
    // Impedance(nfreqs, f_max, oclh): zeros (src/Z/Impedance.cpp:48-57)
    Impedance(size_t nfreqs, frequency_t f_max, oclhptr_t oclh = nullptr)
        : dev_(resolve(oclh))
    {
        check(vfps_impedance_create(dev_->handle(), (uint32_t)nfreqs, f_max, &h_));
    }
    // Impedance(z, f_max, oclh) (src/Z/Impedance.cpp:35-46)
    Impedance(const std::vector<impedance_t>& z, frequency_t f_max,
              oclhptr_t oclh = nullptr)
        : dev_(resolve(oclh))
    {
        std::vector<vfps_cplx> c(z.size());
        for (size_t i = 0; i < z.size(); i++) {
            c[i] = {z[i].real(), z[i].imag()};
        }
        check(vfps_impedance_create_table(dev_->handle(), c.data(),
                                          (uint32_t)z.size(), f_max, &h_));
    }
    Impedance(const Impedance& other) : dev_(other.dev_)
    {
        check(vfps_impedance_create(dev_->handle(), (uint32_t)other.nFreqs(),
                                    0.0f, &h_));
        check(vfps_impedance_add(h_, other.h_));
    }
    virtual ~Impedance() { vfps_impedance_destroy(h_); }

    // operator+= (src/Z/Impedance.cpp:77-92)
    Impedance& operator+=(const Impedance& rhs)
    {
        check(vfps_impedance_add(h_, rhs.h_));
        return *this;
    }

    size_t nFreqs() const
    {
        uint32_t n = 0;
        check(vfps_impedance_nfreqs(h_, &n));
        return n;
    }
    // data() / impedance(): the table, copied from the device
    const std::vector<impedance_t>& impedance() const
    {
        std::vector<vfps_cplx> c(nFreqs());
        check(vfps_impedance_download(h_, c.data()));
        host_.resize(c.size());
        for (size_t i = 0; i < c.size(); i++) {
            host_[i] = impedance_t(c[i].re, c[i].im);
        }
        return host_;
    }
    const impedance_t* data() const { return impedance().data(); }
    impedance_t operator[](size_t n) const { return impedance().at(n); }
    vfps_impedance* handle() const { return h_; }
    const oclhptr_t& device() const { return dev_; }

protected:
    oclhptr_t dev_;
    vfps_impedance* h_ = nullptr;
    mutable std::vector<impedance_t> host_;
};

// FreeSpaceCSR(n, f_rev, f_max) (src/Z/FreeSpaceCSR.cpp:19-42)
class FreeSpaceCSR : public Impedance {
public:
    FreeSpaceCSR(size_t n, frequency_t f_rev, frequency_t f_max,
                 oclhptr_t oclh = nullptr)
        : Impedance(n, f_max, oclh)
    {
        check(vfps_impedance_add_freespace_csr(h_, f_rev, f_max));
    }
};

// ParallelPlatesCSR(nfreqs, f0, f_max, g) (src/Z/ParallelPlatesCSR.cpp:19-69)
class ParallelPlatesCSR : public Impedance {
public:
    ParallelPlatesCSR(size_t nfreqs, frequency_t f0, frequency_t f_max,
                      double g, oclhptr_t oclh = nullptr)
        : Impedance(nfreqs, f_max, oclh)
    {
        check(vfps_impedance_add_parallel_plates(h_, f0, f_max, g));
    }
};

// ResistiveWall(n, f0, f_max, L, s, xi, b) (src/Z/ResistiveWall.cpp:25-55)
class ResistiveWall : public Impedance {
public:
    ResistiveWall(size_t n, frequency_t f0, frequency_t f_max, double L,
                  double s, double xi, double b, oclhptr_t oclh = nullptr)
        : Impedance(n, f_max, oclh)
    {
        check(vfps_impedance_add_resistive_wall(h_, f0, f_max, L, s, xi, b));
    }
};

// ConstImpedance(n, f_max, Z) (src/Z/ConstImpedance.cpp:11-27)
class ConstImpedance : public Impedance {
public:
    ConstImpedance(size_t n, frequency_t f_max, impedance_t Z,
                   oclhptr_t oclh = nullptr)
        : Impedance(n, f_max, oclh)
    {
        check(vfps_impedance_add_const(h_, Z.real(), Z.imag()));
    }
};

// CollimatorImpedance(n, f_max, outer, inner) (src/Z/CollimatorImpedance.cpp)
class CollimatorImpedance : public Impedance {
public:
    CollimatorImpedance(size_t n, frequency_t f_max, double outer,
                        double inner, oclhptr_t oclh = nullptr)
        : Impedance(n, f_max, oclh)
    {
        check(vfps_impedance_add_collimator(h_, outer, inner));
    }
};

// makeImpedance (src/Z/ImpedanceFactory.cpp:22-84): the sum of the selected
// contributions, nullptr if none is selected; `table` stands for the
// impedance file, read by the caller (nullptr: none)
inline std::unique_ptr<Impedance> makeImpedance(
    size_t nfreqs, oclhptr_t oclh, frequency_t fmax, double R_bend,
    double frev, double gap, bool use_csr = true, double s = 0, double xi = 0,
    double inner_coll_radius = 0,
    const std::vector<impedance_t>* table = nullptr)
{
    const double c = 299792458.0;
    const double f0 = c / (2 * 3.14159265358979323846 * R_bend);
    std::unique_ptr<Impedance> rv(new Impedance(nfreqs, fmax, oclh));
    bool changed = false;
    if (gap != 0) {
        if (use_csr) {
            changed = true;
            if (gap > 0) {
                check(vfps_impedance_add_parallel_plates(rv->handle(),
                                                         (frequency_t)f0,
                                                         fmax, gap));
            } else {
                check(vfps_impedance_add_freespace_csr(rv->handle(),
                                                       (frequency_t)f0, fmax));
            }
        }
        const double radius = gap > 0 ? gap / 2 : -gap / 2;
        if (s > 0 && xi >= -1) {
            changed = true;
            check(vfps_impedance_add_resistive_wall(rv->handle(),
                                                    (frequency_t)frev, fmax,
                                                    c / frev, s, xi, radius));
        }
        if (inner_coll_radius > 0 && inner_coll_radius < radius) {
            changed = true;
            check(vfps_impedance_add_collimator(rv->handle(), radius,
                                                inner_coll_radius));
        }
    }
    if (table) {
        changed = true;
        *rv += Impedance(*table, fmax, oclh);
    }
    if (!changed) {
        rv.reset();
    }
    return rv;
}

class ElectricField;

// ---- SourceMaps (inc/SM/SourceMap.hpp:53-87) ----------------------------------
class SourceMap {
public:
    enum InterpolationType : uint_fast8_t {
        none = 1, linear = 2, quadratic = 3, cubic = 4
    };

    SourceMap(const SourceMap&) = delete;
    SourceMap& operator=(const SourceMap&) = delete;
    virtual ~SourceMap() { vfps_map_destroy(h_); }

    // apply (src/SM/SourceMap.cpp:75-109)
    virtual void apply() { check(vfps_map_apply(h_)); }
    vfps_map* handle() const { return h_; }

protected:
    SourceMap(std::shared_ptr<PhaseSpace> in, std::shared_ptr<PhaseSpace> out)
        : in_(std::move(in)), out_(std::move(out)) {}

    std::shared_ptr<PhaseSpace> in_, out_;
    vfps_map* h_ = nullptr;
};

class KickMap : public SourceMap {
public:
    enum class Axis : uint_fast8_t { x = 0, y = 1 };

    // KickMap(in, out, it, interpol_clamp, kd, oclh) (src/SM/KickMap.cpp:25-192)
    KickMap(std::shared_ptr<PhaseSpace> in, std::shared_ptr<PhaseSpace> out,
            InterpolationType it, bool interpol_clamp, Axis kd, oclhptr_t)
        : SourceMap(std::move(in), std::move(out))
    {
        check(vfps_map_create_kick(in_->handle(), out_->handle(), it,
                                   interpol_clamp ? 1 : 0, (int)kd, &h_));
    }

    // swapOffset (inc/SM/KickMap.hpp:51-55): the offsets become the map's,
    // the caller receives the previous ones
    void swapOffset(std::vector<meshaxis_t>& offset)
    {
        std::vector<meshaxis_t> prev = force();
        check(vfps_map_set_offset(h_, offset.data(), offset.size()));
        offset.swap(prev);
    }

    // getForce (inc/SM/KickMap.hpp:47-49)
    const meshaxis_t* getForce()
    {
        force_ = force();
        return force_.data();
    }

protected:
    KickMap(std::shared_ptr<PhaseSpace> in, std::shared_ptr<PhaseSpace> out)
        : SourceMap(std::move(in), std::move(out)) {}

    std::vector<meshaxis_t> force()
    {
        const size_t n = (size_t)in_->nMeshCells(0) * in_->nBunches();
        std::vector<meshaxis_t> f(n);
        check(vfps_map_get_offset(h_, f.data(), f.size()));
        return f;
    }

    std::vector<meshaxis_t> force_;
};

// DriftMap(in, out, slip, E0, it, clamp, oclh) (src/SM/DriftMap.cpp:11-38)
class DriftMap : public KickMap {
public:
    DriftMap(std::shared_ptr<PhaseSpace> in, std::shared_ptr<PhaseSpace> out,
             const std::vector<meshaxis_t>& slip, meshaxis_t E0,
             InterpolationType it, bool interpol_clamp, oclhptr_t)
        : KickMap(std::move(in), std::move(out))
    {
        check(vfps_map_create_drift(in_->handle(), out_->handle(), slip.data(),
                                    (uint32_t)slip.size(), E0, it,
                                    interpol_clamp ? 1 : 0, &h_));
    }
};

// RFKickMap (src/SM/RFKickMap.cpp:15-57)
class RFKickMap : public KickMap {
public:
    // linear
    RFKickMap(std::shared_ptr<PhaseSpace> in, std::shared_ptr<PhaseSpace> out,
              meshaxis_t angle, frequency_t f_RF, InterpolationType it,
              bool interpol_clamp, oclhptr_t)
        : KickMap(std::move(in), std::move(out))
    {
        check(vfps_map_create_rfkick_linear(in_->handle(), out_->handle(),
                                            angle, f_RF, it,
                                            interpol_clamp ? 1 : 0, &h_));
    }
    // sinusoidal
    RFKickMap(std::shared_ptr<PhaseSpace> in, std::shared_ptr<PhaseSpace> out,
              timeaxis_t revolutionpart, meshaxis_t V_RF, frequency_t f_RF,
              meshaxis_t V0, InterpolationType it, bool interpol_clamp,
              oclhptr_t)
        : KickMap(std::move(in), std::move(out))
    {
        check(vfps_map_create_rfkick(in_->handle(), out_->handle(),
                                     revolutionpart, V_RF, f_RF, V0, it,
                                     interpol_clamp ? 1 : 0, &h_));
    }
};

// DynamicRFKickMap (src/SM/DynamicRFKickMap.cpp:15-67): the reference's two
// constructors plus the noise seed, which the reference draws from
// std::random_device.  getPastModulation() returns the (phase, amplitude)
// pairs applied since the last call.
class DynamicRFKickMap : public KickMap {
public:
    // linear RF
    DynamicRFKickMap(std::shared_ptr<PhaseSpace> in,
                     std::shared_ptr<PhaseSpace> out, meshindex_t, meshindex_t,
                     meshaxis_t angle, double revolutionpart, double f_RF,
                     meshaxis_t phasespread, meshaxis_t amplspread,
                     meshaxis_t modampl, double modtimeincrement,
                     uint32_t steps, InterpolationType it, bool interpol_clamp,
                     oclhptr_t, uint32_t seed = 0)
        : KickMap(std::move(in), std::move(out))
    {
        check(vfps_map_create_dynrfkick_linear(
            in_->handle(), out_->handle(), angle, revolutionpart,
            (float)f_RF, phasespread, amplspread, modampl, modtimeincrement,
            steps, seed, it, interpol_clamp ? 1 : 0, &h_));
    }
    // sinusoidal RF
    DynamicRFKickMap(std::shared_ptr<PhaseSpace> in,
                     std::shared_ptr<PhaseSpace> out, meshindex_t, meshindex_t,
                     double revolutionpart, double V_RF, double f_RF, double V0,
                     meshaxis_t phasespread, meshaxis_t amplspread,
                     meshaxis_t modampl, double modtimeincrement,
                     uint32_t steps, InterpolationType it, bool interpol_clamp,
                     oclhptr_t, uint32_t seed = 0)
        : KickMap(std::move(in), std::move(out))
    {
        check(vfps_map_create_dynrfkick(
            in_->handle(), out_->handle(), revolutionpart, (float)V_RF,
            (float)f_RF, (float)V0, phasespread, amplspread, modampl,
            modtimeincrement, steps, seed, it, interpol_clamp ? 1 : 0, &h_));
    }

    std::vector<std::array<meshaxis_t, 2>> getPastModulation()
    {
        uint32_t applied = 0;
        check(vfps_map_get_modulation(h_, 0, 0, nullptr, &applied));
        std::vector<std::array<meshaxis_t, 2>> rv(applied - read_);
        if (!rv.empty()) {
            check(vfps_map_get_modulation(h_, read_, applied - read_,
                                          rv[0].data(), nullptr));
        }
        read_ = applied;
        return rv;
    }

private:
    uint32_t read_ = 0;
};

// RotationMap(in, out, xsize, ysize, angle, it, clamped, rotmapsize, oclh)
// (src/SM/RotationMap.cpp:12-95)
class RotationMap : public SourceMap {
public:
    RotationMap(std::shared_ptr<PhaseSpace> in, std::shared_ptr<PhaseSpace> out,
                meshindex_t, meshindex_t, meshaxis_t angle, InterpolationType it,
                bool interpol_clamped, meshindex_t rotmapsize, oclhptr_t)
        : SourceMap(std::move(in), std::move(out))
    {
        check(vfps_map_create_rotation(in_->handle(), out_->handle(), angle, it,
                                       interpol_clamped ? 1 : 0, rotmapsize,
                                       &h_));
    }
};

// FokkerPlanckMap(in, out, xsize, ysize, fptype, fptrack, e1, dt, oclh)
// (src/SM/FokkerPlanckMap.cpp:11-118)
class FokkerPlanckMap : public SourceMap {
public:
    enum class FPType : uint_fast8_t {
        none = 0, damping_only = 1, diffusion_only = 2, full = 3
    };
    enum class FPTracking : uint_fast8_t {
        none = 0, approximation1 = 1, approximation2 = 2, stochastic = 3
    };
    // two_sided and cubic as in the reference; five_point is this
    // library's extension (exact on quartics)
    enum DerivationType : uint_fast8_t {
        two_sided = 3, cubic = 4, five_point = 5
    };

    FokkerPlanckMap(std::shared_ptr<PhaseSpace> in,
                    std::shared_ptr<PhaseSpace> out, meshindex_t, meshindex_t,
                    FPType fptype, FPTracking fptrack, timeaxis_t e1,
                    DerivationType dt, oclhptr_t)
        : SourceMap(std::move(in), std::move(out))
    {
        check(vfps_map_create_fokkerplanck(in_->handle(), out_->handle(),
                                           (int)fptype, (int)fptrack, e1,
                                           (int)dt, &h_));
    }
};

// Identity (inc/SM/Identity.hpp:14-61)
class Identity : public SourceMap {
public:
    Identity(std::shared_ptr<PhaseSpace> in, std::shared_ptr<PhaseSpace> out,
             oclhptr_t)
        : SourceMap(std::move(in), std::move(out))
    {
        check(vfps_map_create_identity(in_->handle(), out_->handle(), &h_));
    }
};

// ---- ElectricField (inc/PS/ElectricField.hpp) -----------------------------
class ElectricField {
public:
    // ElectricField(ps, impedance, buckets, spacing_bins, oclh, f_rev,
    // revolutionpart, wakescaling) (src/PS/ElectricField.cpp:14-101)
    ElectricField(std::shared_ptr<PhaseSpace> ps,
                  std::shared_ptr<Impedance> impedance,
                  const std::vector<uint32_t>& bucketnumbers,
                  meshindex_t spacing_bins, oclhptr_t, double f_rev,
                  meshaxis_t revolutionpart = 1, meshaxis_t wakescaling = 0.0f)
        : ps_(std::move(ps)), z_(std::move(impedance))
    {
        check(vfps_field_create(ps_->handle(), z_->handle(),
                                bucketnumbers.data(),
                                (uint32_t)bucketnumbers.size(), spacing_bins,
                                f_rev, revolutionpart, wakescaling, &h_));
    }

    // the wake-potential constructor (inc/PS/ElectricField.hpp:81-103)
    ElectricField(std::shared_ptr<PhaseSpace> ps,
                  std::shared_ptr<Impedance> impedance,
                  const std::vector<uint32_t>& bucketnumbers,
                  meshindex_t spacing_bins, oclhptr_t oclh, double f_rev,
                  meshaxis_t revolutionpart, double Ib, double E0,
                  double sigma_delta, double dt)
        : ElectricField(ps, impedance, bucketnumbers, spacing_bins, oclh, f_rev,
                        revolutionpart,
                        vfps_field_wakescaling(Ib, dt, ps->getScale(0),
                                               ps->getDelta(1), sigma_delta,
                                               E0)) {}

    ~ElectricField() { vfps_field_destroy(h_); }
    ElectricField(const ElectricField&) = delete;
    ElectricField& operator=(const ElectricField&) = delete;

    size_t getNMax() const { return z_->nFreqs(); }

    // padBunchProfiles (src/PS/ElectricField.cpp:255-263)
    void padBunchProfiles() { check(vfps_field_pad(h_)); }
    const float* getPaddedBunchProfiles()
    {
        padded_.resize(getNMax());
        check(vfps_field_get_padded(h_, padded_.data()));
        return padded_.data();
    }

    // wakePotential (src/PS/ElectricField.cpp:186-253): [nb][n]
    meshaxis_t* wakePotential()
    {
        wake_.resize((size_t)ps_->nBunches() * ps_->nMeshCells(0));
        check(vfps_field_wake_potential(h_, wake_.data()));
        return wake_.data();
    }

    // updateCSR (src/PS/ElectricField.cpp:133-184): the spectrum
    const csrpower_t* updateCSR(frequency_t cutoff_frequency)
    {
        spectrum_.resize((size_t)ps_->nBunches() * getNMax());
        power_.resize(ps_->nBunches());
        check(vfps_field_update_csr(h_, cutoff_frequency, power_.data(),
                                    spectrum_.data()));
        return spectrum_.data();
    }
    const csrpower_t* getCSRPower() const { return power_.data(); }
    const csrpower_t* getCSRSpectrum() const { return spectrum_.data(); }

    meshaxis_t getWakeScaling() const
    {
        float s = 0;
        check(vfps_field_get_wakescaling(h_, &s));
        return s;
    }
    vfps_field* handle() const { return h_; }

private:
    std::shared_ptr<PhaseSpace> ps_;
    std::shared_ptr<Impedance> z_;
    vfps_field* h_ = nullptr;
    std::vector<float> padded_, wake_, spectrum_, power_;
};

// WakePotentialMap(in, out, field, it, clamp, oclh)
// (src/SM/WakePotentialMap.cpp:11-50)
class WakePotentialMap : public KickMap {
public:
    WakePotentialMap(std::shared_ptr<PhaseSpace> in,
                     std::shared_ptr<PhaseSpace> out, ElectricField* field,
                     InterpolationType it, bool interpol_clamp, oclhptr_t)
        : KickMap(std::move(in), std::move(out))
    {
        check(vfps_map_create_wakepotential(in_->handle(), out_->handle(),
                                            field->handle(), it,
                                            interpol_clamp ? 1 : 0, &h_));
    }
    // update (src/SM/WakePotentialMap.cpp:34-50)
    void update() { check(vfps_map_update(h_)); }
};

}  // namespace inovesa_amd

#endif
