// cuda_backend_hip.cpp -- the reference's CudaDemSolver on top of libdemhip.so.
//
// Drop-in replacement for the reference's src/cuda_backend_cuda.cu and
// src/cuda/*: it implements every member declared in the reference's
// include/cuda_backend.h:17-75 by converting the host's PARTICLE / BODYSET /
// FORCE objects to the plain arrays of include/dem_hip.h and rethrowing C-ABI
// errors as std::runtime_error with the library's message.  Compiled against
// the reference's own headers (-I<reference>/include); see INTEGRATION.md.
//
// With DEM_HIP_ADAPTER_FREE_FUNCTIONS_ONLY only CudaBackendAvailable and
// CudaDeviceDescription are built (for a host that runs its CPU integrators).
#include "cuda_backend.h"
#include "dem_hip.h"

#include <stdexcept>
#include <string>

bool CudaBackendAvailable() { return dem_backend_available() != 0; }
const char* CudaDeviceDescription() { return dem_device_description(); }

#ifndef DEM_HIP_ADAPTER_FREE_FUNCTIONS_ONLY

#include "force.h"
#include "multibody.h"
#include "particle.h"
#include "simulation.h"

#include <vector>

namespace {

void check(int rc)
{
    if (rc != DEM_OK) throw std::runtime_error(dem_last_error());
}

}  // namespace

class CudaDemSolver::Impl {
public:
    ~Impl()
    {
        if (solver) dem_destroy(solver);
    }
    dem_solver* solver = nullptr;
    int n = 0;
    int nb = 0;
    std::vector<double> pose;
    std::vector<double> loads;
};

CudaDemSolver::CudaDemSolver() : impl_(new Impl)
{
    check(dem_create(&impl_->solver));
}

CudaDemSolver::~CudaDemSolver() = default;

namespace {

void fillPose(const BODYSET& b, std::vector<double>& pose)
{
    pose.resize(static_cast<std::size_t>(b.Num) * 13);
    for (int i = 0; i < b.Num; ++i) {
        const BODY& x = b.body[i];
        double* q = pose.data() + 13 * i;
        q[0] = x.MassCenter[0]; q[1] = x.MassCenter[1]; q[2] = x.MassCenter[2];
        q[3] = x.orien[0]; q[4] = x.orien[1]; q[5] = x.orien[2]; q[6] = x.orien[3];
        q[7] = x.Vel[0]; q[8] = x.Vel[1]; q[9] = x.Vel[2];
        q[10] = x.AngularVel[0]; q[11] = x.AngularVel[1]; q[12] = x.AngularVel[2];
    }
}

dem_mech toMech(const MECH& m)
{
    dem_mech r;
    r.epsS = m.epsS; r.mu = m.mu; r.epsN = m.epsN; r.cohesion = m.c; r.beta = m.beta;
    r.muT = m.mu_T; r.muR = m.mu_R; r.kN = m.kN;
    return r;
}

}  // namespace

void CudaDemSolver::initialize(PARTICLE& p, BODYSET* bodies, const CONTROL& control,
                               const MECH& particleMech, const MECH& bodyMech,
                               const MECH& wallMech, const ODE& ode)
{
    const int n = p.Num;
    std::vector<double> inertia(3 * n), pos(3 * n), vel(3 * n), ang(3 * n), quat(4 * n);
    for (int i = 0; i < n; ++i) {
        for (int a = 0; a < 3; ++a) {
            inertia[3 * i + a] = p.Inertia[i][a];
            pos[3 * i + a] = p.Pos[i][a];
            vel[3 * i + a] = p.Vel[i][a];
            ang[3 * i + a] = p.AngSpd[i][a];
        }
        for (int a = 0; a < 4; ++a) quat[4 * i + a] = p.Quat[i][a];
    }
    dem_particles_view pv{n, p.Number, p.Status, p.Mass, p.Radius, inertia.data(), pos.data(),
                          vel.data(), ang.data(), quat.data()};
    std::vector<double> cpb(3 * static_cast<std::size_t>(p.ComponentNum));
    for (int c = 0; c < p.ComponentNum; ++c)
        for (int a = 0; a < 3; ++a) cpb[3 * c + a] = p.ComponentPosBody[c][a];
    dem_components_view cv{p.ComponentNum, p.ComponentId, p.ComponentOwner, p.ComponentRadius,
                           cpb.data()};
    // walls only when enabled (cuda_backend_cuda.cu:336-339)
    std::vector<double> wo, wn;
    dem_walls_view wv{0, nullptr, nullptr};
    if (control.Wall_flag && control.wall.number > 0) {
        const int w = control.wall.number;
        wo.resize(3 * w);
        wn.resize(3 * w);
        for (int i = 0; i < w; ++i)
            for (int a = 0; a < 3; ++a) {
                wo[3 * i + a] = control.wall.Orig[i][a];
                wn[3 * i + a] = control.wall.N[i][a];
            }
        wv = dem_walls_view{w, wo.data(), wn.data()};
    }
    std::vector<double> bmass, tris;
    std::vector<int> tcount;
    dem_bodies_view bv{0, nullptr, nullptr, nullptr, nullptr};
    if (bodies && bodies->Num > 0) {
        const int nb = bodies->Num;
        fillPose(*bodies, impl_->pose);
        for (int b = 0; b < nb; ++b) {
            const BODY& x = bodies->body[b];
            if (x.Size < 0) throw std::runtime_error("body triangle count cannot be negative");
            bmass.push_back(x.Mass);
            tcount.push_back(x.Size);
            for (int k = 0; k < x.Size; ++k) {
                const STLTetMesh& t = x.Shape[k];
                const vector3d* v[5] = {&t.v1, &t.v2, &t.v3, &t.v4, &t.normal};
                for (const vector3d* e : v)
                    for (int a = 0; a < 3; ++a) tris.push_back((*e)[a]);
            }
        }
        bv = dem_bodies_view{nb, bmass.data(), impl_->pose.data(), tcount.data(), tris.data()};
        impl_->nb = nb;
        impl_->loads.assign(6 * static_cast<std::size_t>(nb), 0.0);
    }
    const dem_mech pp = toMech(particleMech), pt = toMech(bodyMech), pw = toMech(wallMech);
    dem_control ctl{};
    ctl.dt = ode.StepSize;
    ctl.gravity[0] = control.g[0]; ctl.gravity[1] = control.g[1]; ctl.gravity[2] = control.g[2];
    ctl.gravity_enabled = control.Grav_flag ? 1 : 0;
    ctl.universal_gravity = control.Universal_gravitation_flag ? 1 : 0;
    ctl.keep_body_particle = control.OutputParticleForce ? 1 : 0;
    ctl.keep_body_detail = control.OutputBodyForce ? 1 : 0;
    ctl.mesh_scale = 0.0;
    check(dem_initialize(impl_->solver, &pv, &cv, &wv, &bv, &pp, &pt, &pw, &ctl));
    impl_->n = n;
}

void CudaDemSolver::computeForces(double, BODYSET* bodies, double** bodyForces,
                                  bool collectDiagnostics)
{
    const double* pose = nullptr;
    if (bodies && bodies->Num > 0) {
        fillPose(*bodies, impl_->pose);
        pose = impl_->pose.data();
    }
    double* loads = (bodyForces && impl_->nb > 0) ? impl_->loads.data() : nullptr;
    check(dem_compute_forces(impl_->solver, pose, loads, collectDiagnostics ? 1 : 0));
    if (loads)
        for (int b = 0; b < impl_->nb; ++b)
            for (int k = 0; k < 6; ++k) bodyForces[b][k] = loads[6 * b + k];
}

void CudaDemSolver::initializeParticleHalfStep(double halfDt)
{
    check(dem_init_half_step(impl_->solver, halfDt));
}

double CudaDemSolver::advanceParticles(double fullDt)
{
    double v = 0.0;
    check(dem_advance_particles(impl_->solver, fullDt, &v));
    return v;
}

void CudaDemSolver::finishParticleStep(double fullDt, double halfDt)
{
    check(dem_finish_particle_step(impl_->solver, fullDt, halfDt));
}

void CudaDemSolver::loadContactHistory(const std::string& filename)
{
    check(dem_load_contact_history(impl_->solver, filename.c_str()));
}

void CudaDemSolver::saveContactHistory(const std::string& filename) const
{
    check(dem_save_contact_history(impl_->solver, filename.c_str()));
}

bool CudaDemSolver::hasRestoredParticleLeapfrogState() const
{
    int v = 0;
    check(dem_has_restored_leapfrog_state(impl_->solver, &v));
    return v != 0;
}

void CudaDemSolver::downloadParticleState(PARTICLE& p) const
{
    const int n = impl_->n;
    if (p.Num != n) throw std::runtime_error("particle count changed during CUDA simulation");
    std::vector<double> pos(3 * n), vel(3 * n), ang(3 * n), quat(4 * n);
    check(dem_download_particle_state(impl_->solver, pos.data(), vel.data(), ang.data(), quat.data()));
    for (int i = 0; i < n; ++i) {
        p.Pos[i] = vector3d(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
        p.Vel[i] = vector3d(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]);
        p.AngSpd[i] = vector3d(ang[3 * i], ang[3 * i + 1], ang[3 * i + 2]);
        p.Quat[i] = vector4d(quat[4 * i], quat[4 * i + 1], quat[4 * i + 2], quat[4 * i + 3]);
    }
}

void CudaDemSolver::downloadForceOutput(FORCE& f, bool particleForces, bool bodyForces) const
{
    const int n = impl_->n, nb = impl_->nb;
    const std::size_t n6 = static_cast<std::size_t>(n) * 6;
    if (particleForces) {
        std::vector<double> c(n6), fi(n6), w(n6), t(n6);
        check(dem_download_force_output(impl_->solver, c.data(), fi.data(), w.data(), t.data(),
                                        nullptr, nullptr));
        for (int p = 0; p < n; ++p)
            for (int k = 0; k < 6; ++k) {
                f.Fc[p][k] = c[6 * p + k];
                f.Ff[p][k] = fi[6 * p + k];
                f.Fw[p][k] = w[6 * p + k];
                f.Fd[p][k] = 0.0;
                f.FR[p][k] = t[6 * p + k];
            }
    }
    if (nb > 0 && (particleForces || bodyForces)) {
        std::vector<double> bp, bd;
        if (particleForces) bp.resize(static_cast<std::size_t>(nb) * n6);
        if (bodyForces) bd.resize(static_cast<std::size_t>(nb) * n * 19);
        check(dem_download_force_output(impl_->solver, nullptr, nullptr, nullptr, nullptr,
                                        particleForces ? bp.data() : nullptr,
                                        bodyForces ? bd.data() : nullptr));
        for (int b = 0; b < nb; ++b)
            for (int p = 0; p < n; ++p) {
                const std::size_t row = static_cast<std::size_t>(b) * n + p;
                if (particleForces)
                    for (int k = 0; k < 6; ++k) f.FI_pt[b][p][k] = bp[row * 6 + k];
                if (bodyForces)
                    for (int k = 0; k < 19; ++k) f.ForceBody[b][p][k] = bd[row * 19 + k];
            }
    }
    const CudaStepStats st = stepStats();
    f.bodyContactCount = st.bodyContacts;
    f.bodyContactMaxDepth = st.bodyMaxDepth;
    f.bodyContactFz = st.bodyContactFz;
}

CudaStepStats CudaDemSolver::stepStats() const
{
    const int nb = impl_->nb;
    dem_step_stats s{};
    std::vector<uint64_t> ri(7 * nb), rv(7 * nb);
    std::vector<double> mi(22 * nb), mv(22 * nb);
    check(dem_get_step_stats(impl_->solver, &s, ri.data(), rv.data(), mi.data(), mv.data()));
    CudaStepStats out;
    out.particleContacts = s.particle_contacts;
    out.bodyContacts = s.body_contacts;
    out.bodyMaxDepth = s.body_max_depth;
    out.bodyContactFz = s.body_contact_fz;
    out.particleHistoryOverflows = s.particle_history_overflows;
    out.bodyHistoryOverflows = s.body_history_overflows;
    out.particleHistoryHighWater = s.particle_history_high_water;
    out.bodyHistoryHighWater = s.body_history_high_water;
    out.triangleGridOverflows = s.triangle_grid_overflows;
    if (nb > 0) {
        out.bodyRegimeInstant.resize(nb);
        out.bodyRegimeInterval.resize(nb);
        out.bodyMicrostateInstant.resize(nb);
        out.bodyMicrostateInterval.resize(nb);
        for (int b = 0; b < nb; ++b) {
            for (int k = 0; k < 7; ++k) {
                out.bodyRegimeInstant[b][k] = ri[7 * b + k];
                out.bodyRegimeInterval[b][k] = rv[7 * b + k];
            }
            for (int k = 0; k < 22; ++k) {
                out.bodyMicrostateInstant[b][k] = mi[22 * b + k];
                out.bodyMicrostateInterval[b][k] = mv[22 * b + k];
            }
        }
        out.bodyRegimeIntervalNativeSteps = s.body_regime_interval_native_steps;
    }
    return out;
}

std::size_t CudaDemSolver::particleGridEntryCount() const
{
    std::size_t v = 0;
    check(dem_particle_grid_entry_count(impl_->solver, &v));
    return v;
}

std::size_t CudaDemSolver::triangleGridEntryCount() const
{
    std::size_t v = 0;
    check(dem_triangle_grid_entry_count(impl_->solver, &v));
    return v;
}

std::size_t CudaDemSolver::deviceMemoryBytes() const
{
    std::size_t v = 0;
    check(dem_device_memory_bytes(impl_->solver, &v));
    return v;
}

#endif  // DEM_HIP_ADAPTER_FREE_FUNCTIONS_ONLY
