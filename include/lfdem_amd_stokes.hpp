// lfdem_amd_stokes.hpp -- header-only C++ drop-in for LF_DEM's StokesSolver
// (StokesSolver.h:31-286) on top of the C ABI of lfdem_amd.h.
//
// Same method names, argument meaning and error behaviour as the reference
// class, so LF_DEM's System can keep its calls unchanged:
//   init -> resetResistanceMatrix -> startNewColumn / addResistanceBlocks
//   -> matrixFillingDone -> resetRHS / addToRHS* / setRHS* -> solve -> solvingIsDone
// DBlock / ODBlock / vec3d are taken as template parameters with LF_DEM's
// member names (MatrixBlocks.h:5-84: col0..col5; vec3d.h: x, y, z), so this
// header needs none of LF_DEM's headers.  Failures throw std::runtime_error,
// as the reference does for an undersized solution vector (StokesSolver.cpp:683-687).
#pragma once
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>
#include "lfdem_amd.h"

namespace lfdem_amd {

inline void check(int rc, const char* what)
{
    if (rc != LFD_OK) throw std::runtime_error(std::string(what) + ": " + lfd_last_error());
}

template <class DB>
inline void pack_dblock(const DB& b, double* o)
{
    for (int k = 0; k < 5; k++) o[k] = b.col0[k];
    for (int k = 0; k < 3; k++) o[5+k] = b.col1[k];
    o[8] = b.col2[0];
    for (int k = 0; k < 3; k++) o[9+k] = b.col3[k];
    for (int k = 0; k < 2; k++) o[12+k] = b.col4[k];
    o[14] = b.col5[0];
}

template <class OD>
inline void pack_odblock(const OD& b, double* o)
{
    for (int k = 0; k < 5; k++) o[k] = b.col0[k];
    for (int k = 0; k < 3; k++) o[5+k] = b.col1[k];
    o[8] = b.col2[0];
    for (int k = 0; k < 5; k++) o[9+k] = b.col3[k];
    for (int k = 0; k < 3; k++) o[14+k] = b.col4[k];
    o[17] = b.col5[0];
}

class StokesSolver {
public:
    StokesSolver() = default;
    StokesSolver(const StokesSolver&) = delete;
    StokesSolver& operator=(const StokesSolver&) = delete;
    ~StokesSolver() { if (h_) lfd_stokes_destroy(h_); }

    // StokesSolver::init (StokesSolver.cpp:77-106)
    void init(int np_total, int np_mobile)
    {
        if (h_) { lfd_stokes_destroy(h_); h_ = nullptr; }
        check(lfd_stokes_create(np_total, np_mobile, &h_), "StokesSolver::init");
        np_ = np_total;
        np_mobile_ = np_mobile;
    }
    void setSolveOptions(const lfd_solve_opts& o) { check(lfd_stokes_set_opts(h_, &o), "setSolveOptions"); }

    // StokesSolver::resetResistanceMatrix (StokesSolver.cpp:474-511); the
    // mobile-fixed / fixed-fixed counts must be 0 (no fixed particles).
    template <class DB>
    void resetResistanceMatrix(int nb_of_interactions_mm, int nb_of_interactions_mf, int nb_of_interactions_ff,
                               const std::vector<DB>& reset_resmat_dblocks, bool matrix_pattern_changed)
    {
        if (nb_of_interactions_mf != 0 || nb_of_interactions_ff != 0)
            throw std::runtime_error("StokesSolver: fixed particles are not supported");
        drag_.resize(15*reset_resmat_dblocks.size());
        for (size_t i = 0; i < reset_resmat_dblocks.size(); i++) pack_dblock(reset_resmat_dblocks[i], &drag_[15*i]);
        check(lfd_stokes_reset(h_, nb_of_interactions_mm, drag_.data(), matrix_pattern_changed), "resetResistanceMatrix");
    }
    void startNewColumn() { check(lfd_stokes_start_new_column(h_), "startNewColumn"); }
    // StokesSolver::addResistanceBlocks (StokesSolver.cpp:121-129)
    template <class DB, class OD>
    void addResistanceBlocks(int i, int j, const std::pair<DB, DB>& DiagBlocks_i_and_j, const OD& ODBlock_ij)
    {
        double a[15], b[15], c[18];
        pack_dblock(DiagBlocks_i_and_j.first, a);
        pack_dblock(DiagBlocks_i_and_j.second, b);
        pack_odblock(ODBlock_ij, c);
        check(lfd_stokes_add_blocks(h_, i, j, a, b, c), "addResistanceBlocks");
    }
    void matrixFillingDone() { check(lfd_stokes_fill_done(h_), "matrixFillingDone"); }

    // RHS (StokesSolver.cpp:513-624)
    void resetRHS() { check(lfd_stokes_reset_rhs(h_), "resetRHS"); }
    void resetRHStorque() { check(lfd_stokes_reset_rhs_torque(h_), "resetRHStorque"); }
    template <class V> void addToRHSForce(int i, const V& f)
    {
        if (i >= np_mobile_) return;                 // fixed particles are skipped (StokesSolver.cpp:531-539)
        double v[3] = {f.x, f.y, f.z};
        check(lfd_stokes_add_rhs_force(h_, i, v), "addToRHSForce");
    }
    template <class V> void addToRHSTorque(int i, const V& t)
    {
        if (i >= np_mobile_) return;
        double v[3] = {t.x, t.y, t.z};
        check(lfd_stokes_add_rhs_torque(h_, i, v), "addToRHSTorque");
    }
    template <class V> void setRHSForce(int i, const V& f)
    {
        if (i >= np_mobile_) return;
        double v[3] = {f.x, f.y, f.z};
        check(lfd_stokes_set_rhs_force(h_, i, v), "setRHSForce");
    }
    template <class V> void setRHSTorque(int i, const V& t)
    {
        if (i >= np_mobile_) return;
        double v[3] = {t.x, t.y, t.z};
        check(lfd_stokes_set_rhs_torque(h_, i, v), "setRHSTorque");
    }
    void addToRHS(const std::vector<double>& force) { check(lfd_stokes_add_rhs(h_, force.data()), "addToRHS"); }
    template <class V> void setRHS(const std::vector<V>& force, const std::vector<V>& torque)
    {
        std::vector<double> ft(6*(size_t)np_);
        for (int i = 0; i < np_mobile_; i++) {
            ft[6*i] = force[i].x; ft[6*i+1] = force[i].y; ft[6*i+2] = force[i].z;
            ft[6*i+3] = torque[i].x; ft[6*i+4] = torque[i].y; ft[6*i+5] = torque[i].z;
        }
        check(lfd_stokes_set_rhs(h_, ft.data()), "setRHS");
    }

    // StokesSolver::solve (StokesSolver.cpp:671-703)
    template <class V> void solve(std::vector<V>& velocity, std::vector<V>& ang_velocity)
    {
        if ((int)velocity.size() < np_mobile_ || (int)ang_velocity.size() < np_mobile_)
            throw std::runtime_error(" StokesSolver::solve: allocated solution vector too small");
        sol_.resize(6*(size_t)np_);
        check(lfd_stokes_solve(h_, sol_.data(), (int)sol_.size(), &stats_), "solve");
        for (int i = 0; i < np_mobile_; i++) {
            velocity[i].x = sol_[6*i]; velocity[i].y = sol_[6*i+1]; velocity[i].z = sol_[6*i+2];
            ang_velocity[i].x = sol_[6*i+3]; ang_velocity[i].y = sol_[6*i+4]; ang_velocity[i].z = sol_[6*i+5];
        }
    }
    void solvingIsDone() { check(lfd_stokes_solving_done(h_), "solvingIsDone"); }
    const lfd_solve_stats& lastSolveStats() const { return stats_; }

private:
    lfd_stokes* h_ = nullptr;
    int np_ = 0, np_mobile_ = 0;
    std::vector<double> drag_, sol_;
    lfd_solve_stats stats_{};
};

}  // namespace lfdem_amd
