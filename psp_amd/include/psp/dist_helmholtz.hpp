// DistHelmholtz<Real> (include/psp/dist_helmholtz.hpp:15-283 of the
// reference) on one MI355X.  Same constructor defaults, Initialize / Solve /
// Finalize semantics and exceptions; SolveWithGMRES is the public name of
// the reference's private InternalSolveWithGMRES path (it forwards to
// Solve).  Real must be double: GMRES runs in complex double; the front
// precision (complex double or complex float) is a constructor option.
#pragma once

#include <complex>
#include <cstdio>
#include <string>
#include <type_traits>
#include <vector>

namespace psp {

enum PanelScheme { CLIQUE_LDL_1D = PSP_HIP_LDL_1D, CLIQUE_LDL_SELINV_1D = PSP_HIP_LDL_SELINV_1D };
enum FrontPrecision { FRONTS_FP64 = PSP_HIP_FRONTS_FP64, FRONTS_FP32 = PSP_HIP_FRONTS_FP32 };

template <typename Real>
class DistHelmholtz {
    static_assert(std::is_same<Real, double>::value, "the MI355X build solves in double");

public:
    typedef Complex<Real> C;

    DistHelmholtz(const Discretization<Real>& disc, mpi::Comm comm, Real damping = 7.0,
                  int numPlanesPerPanel = 4, int cutoff = 12,
                  FrontPrecision precision = FRONTS_FP64)
        : disc_(disc), comm_(comm)
    {
        psp_hip_desc d{};
        d.omega = disc.omega;
        d.nx = disc.nx;
        d.ny = disc.ny;
        d.nz = disc.nz;
        d.wx = disc.wx;
        d.wy = disc.wy;
        d.wz = disc.wz;
        d.frontBC = disc.frontBC;
        d.rightBC = disc.rightBC;
        d.backBC = disc.backBC;
        d.leftBC = disc.leftBC;
        d.topBC = disc.topBC;
        d.bx = disc.bx;
        d.by = disc.by;
        d.bz = disc.bz;
        d.sigmax = disc.sigmax;
        d.sigmay = disc.sigmay;
        d.sigmaz = disc.sigmaz;
        d.damping = damping;
        d.planesPerPanel = numPlanesPerPanel;
        d.cutoff = cutoff;
        d.precision = precision;
        const int st = psp_hip_create(&d, &ctx_);
        if (st == PSP_HIP_ERR_SHALLOW)
            LogicError("The domain is very shallow. Please run a sparse-direct factorization "
                       "instead.");
        if (st != PSP_HIP_OK) LogicError("invalid discretization (status ", st, ")");
    }

    ~DistHelmholtz()
    {
        if (ctx_) psp_hip_destroy(ctx_);
    }
    DistHelmholtz(const DistHelmholtz&) = delete;
    DistHelmholtz& operator=(const DistHelmholtz&) = delete;

    // Build the sparse matrix and the preconditioner
    // (initialize_finalize.hpp:12-377).
    void Initialize(const DistUniformGrid<Real>& velocity,
                    PanelScheme panelScheme = CLIQUE_LDL_SELINV_1D)
    {
        if (disc_.nx != velocity.XSize() || disc_.ny != velocity.YSize() ||
            disc_.nz != velocity.ZSize())
            LogicError("Velocity grid is incorrect");
        if (velocity.NumScalars() != 1) LogicError("Velocity grid should have one entry per point");
        if (velocity.Order() == XYZ) {
            ThrowOnStatus(psp_hip_factor(ctx_, velocity.LockedBuffer(), panelScheme,
                                         &factorReport_),
                          ctx_, "Initialize");
        } else {
            const std::vector<Real> nat = velocity.Natural();
            ThrowOnStatus(psp_hip_factor(ctx_, nat.data(), panelScheme, &factorReport_), ctx_,
                          "Initialize");
        }
    }

    // Solves an O(1) set of linear systems with the sweeping preconditioner
    // (solve.hpp:12-56); B is overwritten with the solutions.
    // With viewIterates every unconverged GMRES iterate is written as the VTK
    // volume "iterates-<it>" (solve.hpp:579-585).
    void Solve(DistUniformGrid<C>& B, int m = 20, Real relTol = 1e-4,
               bool viewIterates = false) const
    {
        if (B.XSize() != disc_.nx || B.YSize() != disc_.ny || B.ZSize() != disc_.nz)
            LogicError("B grid is incorrect");
        struct View {
            DistUniformGrid<C>* grid;
            static void Dump(void* user, int it, const void* x, int)
            {
                DistUniformGrid<C>& g = *((View*)user)->grid;
                g.SetNatural((const C*)x);
                g.WriteVolume("iterates-" + std::to_string(it));
            }
        } view{&B};
        psp_hip_set_iterate_callback(ctx_, viewIterates ? &View::Dump : nullptr, &view);
        int st;
        if (B.Order() == XYZ) {
            st = psp_hip_solve(ctx_, B.Buffer(), B.NumScalars(), m, relTol, 0, 0, &solveReport_);
        } else {
            std::vector<C> nat = B.Natural();
            st = psp_hip_solve(ctx_, nat.data(), B.NumScalars(), m, relTol, 0, 0, &solveReport_);
            B.SetNatural(nat.data());
        }
        psp_hip_set_iterate_callback(ctx_, nullptr, nullptr);
        ThrowOnStatus(st, ctx_, "Solve");
    }

    void SolveWithGMRES(DistUniformGrid<C>& B, int m = 20, Real relTol = 1e-4,
                        bool viewIterates = false) const
    {
        Solve(B, m, relTol, viewIterates);
    }

    // Destroy the sparse matrix and the preconditioner
    // (initialize_finalize.hpp:379-396); the object stays usable and a later
    // Initialize rebuilds them.
    void Finalize() { ThrowOnStatus(psp_hip_finalize(ctx_), ctx_, "Finalize"); }

    const psp_hip_factor_report& FactorReport() const { return factorReport_; }
    const psp_hip_solve_report& SolveReport() const { return solveReport_; }

private:
    Discretization<Real> disc_;
    mpi::Comm comm_;
    psp_hip_ctx* ctx_ = nullptr;
    psp_hip_factor_report factorReport_{};
    mutable psp_hip_solve_report solveReport_{};
};

}  // namespace psp
