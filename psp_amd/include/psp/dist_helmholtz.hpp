// DistHelmholtz<Real> (include/psp/dist_helmholtz.hpp:15-283 of the
// reference) on one MI355X.  Same constructor defaults, Initialize / Solve /
// Finalize semantics and exceptions; SolveWithGMRES is the public name of
// the reference's private InternalSolveWithGMRES path (it forwards to
// Solve).  Real must be double: GMRES runs in complex double; the front
// precision (complex double or complex float) is a constructor option.
#pragma once

#include <complex>
#include <cstdio>
#include <type_traits>

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

    ~DistHelmholtz() { Finalize(); }
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
        ThrowOnStatus(psp_hip_factor(ctx_, velocity.LockedBuffer(), panelScheme, &factorReport_),
                      ctx_, "Initialize");
    }

    // Solves an O(1) set of linear systems with the sweeping preconditioner
    // (solve.hpp:12-56); B is overwritten with the solutions.
    void Solve(DistUniformGrid<C>& B, int m = 20, Real relTol = 1e-4,
               bool viewIterates = false) const
    {
        (void)viewIterates;  // VTK iterate dumps are not written
        const int st = psp_hip_solve(ctx_, B.Buffer(), B.NumScalars(), m, relTol, 0, 0,
                                     &solveReport_);
        ThrowOnStatus(st, ctx_, "Solve");
    }

    void SolveWithGMRES(DistUniformGrid<C>& B, int m = 20, Real relTol = 1e-4,
                        bool viewIterates = false) const
    {
        Solve(B, m, relTol, viewIterates);
    }

    // Destroy the sparse matrix and the preconditioner
    // (initialize_finalize.hpp:379-396).
    void Finalize()
    {
        if (ctx_) psp_hip_destroy(ctx_);
        ctx_ = nullptr;
    }

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
