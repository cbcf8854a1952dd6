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
//   m_TH (SphBase.h:546-583)                    m_TH (sphgpu_params)
//
// Errors: the reference asserts; these classes throw SphGpuError carrying
// the library's status code and message.  Like SphBase, an object is not
// thread-safe.
#pragma once

#include "../../include/sphgpu.h"

#include <cstdint>
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

protected:
    virtual void sphStep() { check(sphgpu_step(h_, &lastStats_)); }

    void check(int rc) {
        if (rc != SPHGPU_OK) throw SphGpuError(rc, h_ ? sphgpu_last_error(h_) : "sphgpu call failed");
    }

private:
    void ensureHandle() {
        if (h_) return;
        int rc = sphgpu_create(&m_TH, &h_);
        if (rc != SPHGPU_OK) throw SphGpuError(rc, "sphgpu_create: invalid parameters");
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
        if (m_TH.solver != SPHGPU_DFSPH) check(sphgpu_get_fluid(h_, SPHGPU_F_PRESSURE, p.data()));
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
        // static solids: positions are the caller's samples; volumes in add order
        size_t k = 0;
        for (size_t s = 0; s < solidXyz_.size(); ++s)
            for (size_t i = 0; i < solidXyz_[s].size() / 3; ++i, ++k) {
                BoundPart b;
                b.position = {solidXyz_[s][3 * i], solidXyz_[s][3 * i + 1], solidXyz_[s][3 * i + 2]};
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
    std::vector<std::vector<FluidPart>> mirror_;
    std::vector<std::vector<BoundPart>> boundMirror_;
};

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

}  // namespace sphgpu_host
