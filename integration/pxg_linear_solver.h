// pxg_linear_solver.h -- a linear-solver back-end for the reference's
// LinearSolver dispatcher (src/LinearSolver/LinearSolver.cpp:30-121), over
// libpxg: the reference's default solver (DefaultSolver.cpp: RCM ordering,
// skyline envelope, LDU without pivoting, substitution) factored on the GPU
// (pxg_solve_set_pattern, pxg_solve_options, pxg_solve).  Same contract as
// DefaultSolver: initSolver takes the CSR pattern and the "use_rcm" /
// "max_profile_gb" parameters, solveLinearSystem returns false (and prints
// the reason) on a zero pivot or any device failure, so NonlinearSolver
// refuses the update (NonlinearSolver.cpp:100-110) as it does for the CPU
// solvers.
#pragma once

#include "LinearSolver/LinearSolverBase.h"

struct pxg_ctx;

class PxgLinearSolver : public LinearSolverBase {
public:
    PxgLinearSolver() = default;
    ~PxgLinearSolver() override;
    void initSolver(SparseMatrix &K, nlohmann::ordered_json &Parameters) override;
    bool solveLinearSystem(SparseMatrix &A, VectorXd &b, VectorXd &x) override;

private:
    bool setPattern(SparseMatrix &A);
    pxg_ctx *m_Ctx = nullptr;
    int m_Rows = 0, m_NNZ = 0;
    bool m_UseRCM = true;
    double m_MaxProfileGB = 16.0;
};
