// SphGpu.h -- header-only C++ host classes over libsphgpu.so that keep the
// reference's solver class names and call pattern (Source/SphBase.h,
// WcSph.h, DFSPH.h, IISph_M.h), so host code written against SphBase
// (a setupScene() override, runOneStep() in the main loop, particle getters
// for drawing and file output) drives the MI355X particle loops unchanged.
//
//   reference                                  here
//   SphBase::setupScene  (SphBase.h:392)        virtual setupScene()
//   SphBase::stepEvent   (SphBase.h:408)        virtual stepEvent()
//   SphBase::sphStep     (SphBase.h:411)        sphStep() -> sphgpu_step
//   SphBase::runOneStep  (SphBase.cpp:348-354)  runOneStep()
//   getFluidParticles(i) (SphBase.h:513-514)    getFluidParticles(i): lazy AoS mirror
//   getBoundaryParticles (SphBase.h:515-516)    getBoundaryParticles(i)
//   SphWithIO::addFluidCuboid (SphWithIO.cpp:149-194)     addFluidCuboid()
//   SphWithIO::makeCubeSampleParts (SphWithIO.cpp:489-522) addBoxWalls()
//   SphWithBullet::addSolid, rigid (SphWithBullet.cpp:101-136) addRigid()
//   SphWithBullet::setSolidTransform (SphWithBullet.h:73-80)  setSolidTransform()
//   SphSolid::rigidForceJoin (SphBase.cpp:131-143)            rigidForceJoin()
//   SphWithBullet::updateSolids (SphWithBullet.cpp:180-224)   virtual updateSolids(),
//       called inside sphStep where the solvers call it (WcSph.cpp:85,
//       IISph_M.cpp:64, DFSPH.cpp:54)
//   m_TH (SphBase.h:546-583)                    m_TH (sphgpu_params)
//   fUtility.h write_fluid_particles (254-303), write_fluid_particles_json
//     (390-427), write_solid_particles (459-484), read_fluid_particles
//     (486-510)                                 free functions of the same names
//
// Errors: the reference asserts; these classes throw SphGpuError carrying
// the library's status code and message.  Like SphBase, an object is not
// thread-safe.
#pragma once

#include "../../include/sphgpu.h"

#include <array>
#include <cstdint>
#include <exception>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace sphgpu_host {

struct vec_t {
    float x = 0, y = 0, z = 0;
};

// The fields of FluidParticle (SphBase.h:87-165) that host code reads.
struct FluidPart {
    vec_t position, velocity;
    float density = 0, presure = 0;   // reference spelling (SphBase.h:109)
    long long id = 0;                 // original id (add order)
};

struct BoundPart {
    vec_t position;
    float volume = 0;                 // Akinci volume (SphBase.h:180)
};

class SphGpuError : public std::runtime_error {
public:
    SphGpuError(int code, const std::string& what) : std::runtime_error(what), code_(code) {}
    int code() const { return code_; }

private:
    int code_;
};

class SphBase {
public:
    sphgpu_params m_TH;

    explicit SphBase(int solver) {
        sphgpu_default_params(&m_TH);
        m_TH.solver = solver;
    }
    virtual ~SphBase() { sphgpu_destroy(h_); }
    SphBase(const SphBase&) = delete;
    SphBase& operator=(const SphBase&) = delete;

    // The scene is built here: addFluidCuboid / addBoxWalls / addFluidParts.
    virtual void setupScene() = 0;
    // Called before each step (emitters in the reference change N here;
    // with this library particles are added before the first step only).
    virtual void stepEvent() {}

    // SphBase::runOneStep: stepEvent -> sphStep -> frameNumber++, time += dt
    void runOneStep() {
        ensureScene();
        stepEvent();
        sphStep();
        ++frameNumber_;
        systemTime_ += lastStats_.dt;
        mirrorValid_ = false;
        boundValid_ = false;
    }

    int getFrameNumber() const { return frameNumber_; }
    double getSystemTime() const { return systemTime_; }
    const sphgpu_stats& lastStats() const { return lastStats_; }
    sphgpu_handle* handle() { return h_; }

    int numFluids() const { return int(phases_.size()); }
    int numSolids() const { return int(solidXyz_.size()); }

    // Fluid particles of phase k, ascending original id (synchronises).
    const std::vector<FluidPart>& getFluidParticles(int k) {
        refreshMirror();
        return mirror_.at(size_t(k));
    }
    const std::vector<BoundPart>& getBoundaryParticles(int s) {
        refreshBoundary();
        return boundMirror_.at(size_t(s));
    }

    // SphFluid(restDensity, viscosity_alpha): returns the phase index
    int addFluid(float restDensity, float viscosityAlpha) {
        ensureHandle();
        int ph = -1;
        check(sphgpu_add_phase(h_, restDensity, viscosityAlpha, &ph));
        phases_.push_back(ph);
        return ph;
    }

    // n particles (xyz AoS) of phase k; velocities zero unless given
    void addFluidParts(int k, const std::vector<float>& xyz, const std::vector<float>* vel = nullptr) {
        ensureHandle();
        long long n = (long long)(xyz.size() / 3);
        long long first = 0;
        check(sphgpu_add_fluid(h_, 0, k, n, xyz.data(), vel ? vel->data() : nullptr, &first));
        for (long long i = 0; i < n; ++i) phaseOfId_.push_back(k);
    }

    // SphWithIO::addFluidCuboid: an nx*ny*nz lattice from `origin`
    void addFluidCuboid(int k, const int n[3], const float origin[3], float jitter = 0.0f,
                        unsigned long long seed = 0) {
        long long cnt = (long long)n[0] * n[1] * n[2];
        std::vector<float> xyz(size_t(3 * cnt));
        sphgpu_scene_lattice(n, origin, m_TH.spacing_r, jitter, seed, 0, xyz.data());
        addFluidParts(k, xyz);
    }

    // A static solid from boundary samples; returns its index
    int addSolid(const std::vector<float>& xyz, float viscosityAlpha) {
        ensureHandle();
        int s = -1;
        check(sphgpu_add_solid(h_, 0, (long long)(xyz.size() / 3), xyz.data(), viscosityAlpha, &s));
        solidXyz_.push_back(xyz);
        return s;
    }

    // SphWithIO::makeCubeSampleParts + addSolid: walls of an axis-aligned box
    int addBoxWalls(const float center[3], const float half[3], bool haveTop, float viscosityAlpha) {
        long long n = sphgpu_scene_box(center, half, m_TH.spacing_r, haveTop ? 1 : 0, nullptr);
        std::vector<float> xyz(size_t(3 * n));
        sphgpu_scene_box(center, half, m_TH.spacing_r, haveTop ? 1 : 0, xyz.data());
        return addSolid(xyz, viscosityAlpha);
    }

    // A dynamic rigid body from samples in body coordinates (innerPositions);
    // placed by setSolidTransform (identity until then).  Returns its index.
    int addRigid(const std::vector<float>& localXyz, float viscosityAlpha) {
        ensureHandle();
        int s = -1;
        check(sphgpu_add_rigid(h_, 0, (long long)(localXyz.size() / 3), localXyz.data(), viscosityAlpha, &s));
        solidXyz_.push_back(localXyz);
        solidPose_.resize(solidXyz_.size(), identity());
        solidPose_[size_t(s)] = identity();
        dynamic_.resize(solidXyz_.size(), false);
        dynamic_[size_t(s)] = true;
        return s;
    }

    // The body's transform (OpenGL column-major 4x4 of its body-frame
    // samples), centre of mass, linear and angular velocity, as the
    // rigid-body step left them.  Takes effect at the next update point.
    void setSolidTransform(int s, const float T[16], const vec_t& com, const vec_t& v, const vec_t& w) {
        ensureHandle();
        const float c[3] = {com.x, com.y, com.z}, lv[3] = {v.x, v.y, v.z}, av[3] = {w.x, w.y, w.z};
        check(sphgpu_set_solid_pose(h_, s, T, c, lv, av));
        for (int q = 0; q < 16; ++q) solidPose_.at(size_t(s))[size_t(q)] = T[q];
        boundValid_ = false;
    }

    // Total force and torque (about the centre of mass) the fluid exerted
    // on dynamic solid s in the last step.
    void rigidForceJoin(int s, double f[3], double t[3]) { check(sphgpu_get_solid_force_torque(h_, s, f, t)); }

    // Frame output at video cadence: capture now (the transfer overlaps the
    // next steps); the write_* functions below write the last capture.
    void snapshotBegin() {
        ensureScene();
        check(sphgpu_snapshot_begin(h_));
    }
    void writePos(const char* file, int idx, unsigned components, const char* logName) {
        ensureScene();
        check(sphgpu_write_pos(h_, file, idx, components, logName));
    }
    void writeJson(const char* file, int idx) {
        ensureScene();
        check(sphgpu_write_json(h_, file, idx));
    }
    void writeSolidPos(const char* file, int idx, const char* logName) {
        ensureScene();
        check(sphgpu_write_solid_pos(h_, file, idx, logName));
    }
    // SphBase::getLogFileName: the name written on line 1 of .pos files
    std::string logFileName = "sph.log";

protected:
    virtual void sphStep() {
        const int rc = sphgpu_step(h_, &lastStats_);
        if (cbError_) {                  // an exception thrown by updateSolids
            std::exception_ptr e = cbError_;
            cbError_ = nullptr;
            std::rethrow_exception(e);
        }
        check(rc);
    }

    // The host rigid-body step (Bullet in the reference): read rigidForceJoin,
    // advance the bodies, setSolidTransform.  Runs inside sphStep.
    virtual void updateSolids() {}

    void check(int rc) {
        if (rc != SPHGPU_OK) throw SphGpuError(rc, h_ ? sphgpu_last_error(h_) : "sphgpu call failed");
    }

private:
    void ensureHandle() {
        if (h_) return;
        int rc = sphgpu_create(&m_TH, &h_);
        if (rc != SPHGPU_OK) throw SphGpuError(rc, "sphgpu_create: invalid parameters");
        check(sphgpu_set_update_solids(h_, &SphBase::updateSolidsThunk, this));
    }
    // no exception may cross the library's C frames
    static int updateSolidsThunk(sphgpu_handle*, void* self) {
        SphBase* me = static_cast<SphBase*>(self);
        try {
            me->updateSolids();
            return 0;
        } catch (...) {
            me->cbError_ = std::current_exception();
            return 1;
        }
    }
    static std::array<float, 16> identity() {
        std::array<float, 16> T{};
        T[0] = T[5] = T[10] = T[15] = 1.0f;
        return T;
    }
    void ensureScene() {
        if (sceneBuilt_) return;
        setupScene();
        ensureHandle();
        check(sphgpu_init(h_));
        sceneBuilt_ = true;
    }
    void refreshMirror() {
        if (mirrorValid_) return;
        ensureScene();
        const size_t n = static_cast<size_t>(sphgpu_num_fluid(h_));
        std::vector<float> pos(3 * n), vel(3 * n), rho(n), p(n);
        std::vector<long long> ids(n);
        check(sphgpu_get_fluid_ids(h_, ids.data()));
        check(sphgpu_get_fluid(h_, SPHGPU_F_POS, pos.data()));
        check(sphgpu_get_fluid(h_, SPHGPU_F_VEL, vel.data()));
        check(sphgpu_get_fluid(h_, SPHGPU_F_RHO, rho.data()));
        if (m_TH.solver != SPHGPU_DFSPH && m_TH.solver != SPHGPU_IISPH_DF)
            check(sphgpu_get_fluid(h_, SPHGPU_F_PRESSURE, p.data()));
        mirror_.assign(phases_.size(), {});
        for (size_t i = 0; i < n; ++i) {
            FluidPart f;
            f.position = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
            f.velocity = {vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]};
            f.density = rho[i];
            f.presure = p[i];
            f.id = ids[i];
            mirror_[size_t(phaseOfId_[size_t(ids[i])])].push_back(f);
        }
        mirrorValid_ = true;
    }
    void refreshBoundary() {
        if (boundValid_) return;
        ensureScene();
        size_t total = 0;
        for (const auto& x : solidXyz_) total += x.size() / 3;
        std::vector<float> vol(total);
        check(sphgpu_get_boundary_volume(h_, vol.data()));
        boundMirror_.assign(solidXyz_.size(), {});
        // static solids: the caller's samples; dynamic ones placed by their
        // last transform (SphBase.cpp:95-128); volumes in add order
        size_t k = 0;
        for (size_t s = 0; s < solidXyz_.size(); ++s)
            for (size_t i = 0; i < solidXyz_[s].size() / 3; ++i, ++k) {
                const float x = solidXyz_[s][3 * i], y = solidXyz_[s][3 * i + 1], z = solidXyz_[s][3 * i + 2];
                BoundPart b;
                b.position = {x, y, z};
                if (s < dynamic_.size() && dynamic_[s]) {
                    const auto& T = solidPose_[s];
                    const float a = T[3] * x + T[7] * y + T[11] * z + T[15];
                    b.position = {(T[0] * x + T[4] * y + T[8] * z + T[12]) / a,
                                  (T[1] * x + T[5] * y + T[9] * z + T[13]) / a,
                                  (T[2] * x + T[6] * y + T[10] * z + T[14]) / a};
                }
                b.volume = vol[k];
                boundMirror_[s].push_back(b);
            }
        boundValid_ = true;
    }

    sphgpu_handle* h_ = nullptr;
    sphgpu_stats lastStats_{};
    bool sceneBuilt_ = false;
    bool mirrorValid_ = false, boundValid_ = false;
    int frameNumber_ = 0;
    double systemTime_ = 0.0;
    std::vector<int> phases_;
    std::vector<int> phaseOfId_;
    std::vector<std::vector<float>> solidXyz_;
    std::vector<std::array<float, 16>> solidPose_;
    std::vector<bool> dynamic_;
    std::exception_ptr cbError_;
    std::vector<std::vector<FluidPart>> mirror_;
    std::vector<std::vector<BoundPart>> boundMirror_;
};

// fUtility.h's writers and reader, same names and argument order
inline void write_fluid_particles(const char* filename, SphBase& sph, int idx = 0) {
    sph.writePos(filename, idx, SPHGPU_POS_POSITION, sph.logFileName.c_str());
}
inline void write_fluid_particles_json(const char* filename, SphBase& sph, int idx = 0) {
    sph.writeJson(filename, idx);
}
inline void write_solid_particles(const char* filename, SphBase& sph, int idx = 0) {
    sph.writeSolidPos(filename, idx, sph.logFileName.c_str());
}
inline void read_fluid_particles(const char* filename, std::vector<float>& data, std::string* maskstr = nullptr) {
    long long n = 0, frame = 0;
    char mask[512];
    int rc = sphgpu_read_pos(filename, nullptr, 0, &n, &frame, mask, int(sizeof(mask)));
    if (rc != SPHGPU_OK) throw SphGpuError(rc, std::string("read_fluid_particles: ") + filename);
    const std::string m(mask);
    long long per = 0;
    const char* names[] = {"position", "velocity", "color", "density", "presure", "acceleration", "dt"};
    const long long sizes[] = {3, 3, 4, 1, 1, 3, 1};
    for (int k = 0; k < 7; ++k)
        if (m.find(names[k]) != std::string::npos) per += sizes[k];
    data.resize(size_t(n * per));
    rc = sphgpu_read_pos(filename, data.data(), (long long)data.size(), &n, &frame, nullptr, 0);
    if (rc != SPHGPU_OK) throw SphGpuError(rc, std::string("read_fluid_particles: ") + filename);
    if (maskstr) *maskstr = m;
}

// WcSph (WcSph.h:23-96): weakly compressible SPH, Tait EOS
class WcSph : public SphBase {
public:
    WcSph() : SphBase(SPHGPU_WCSPH) {}
};

// DFSPH (DFSPH.h:23-80): divergence-free + constant-density solvers
class DFSPH : public SphBase {
public:
    DFSPH() : SphBase(SPHGPU_DFSPH) {}
};

// IISph_M (IISph_M.h:23-95): implicit incompressible SPH, relaxed Jacobi
class IISph_M : public SphBase {
public:
    IISph_M() : SphBase(SPHGPU_IISPH) {}
};

// IISph (IISph.h:10-100) with its DF_* solvers (IISph.cpp:100-288, 816-1217):
// per-particle mass rho0 beta d^3 (setFluidBeta), m_TH.adjustDensity and
// m_TH.boundaryDensity (adjust_density, boundary_density), and the
// reference's tolerances as defaults
class IISph : public SphBase {
public:
    IISph() : SphBase(SPHGPU_IISPH_DF) {
        m_TH.density_tol = 1e-4f;          // IISph.cpp:826
        m_TH.min_iter_density = 3;         // IISph.cpp:829
        m_TH.max_iter_density = 100;       // IISph.cpp:828
        m_TH.divergence_tol = 1e-3f;       // IISph.cpp:1047 (on mean(Drho/Dt) dt)
        m_TH.min_iter_div = 1;             // IISph.cpp:1050
        m_TH.max_iter_div = 30;            // IISph.cpp:1049
    }
    // FluidParticle::beta of the particles [first, first + n) in add order,
    // after they were added and before the first step
    void setFluidBeta(long long first, const std::vector<float>& beta) {
        if (!handle()) throw SphGpuError(SPHGPU_E_STATE, "setFluidBeta: no fluid added yet");
        check(sphgpu_set_fluid_beta(handle(), first, (long long)beta.size(), beta.data()));
    }
};

}  // namespace sphgpu_host
