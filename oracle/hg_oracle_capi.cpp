// hg_oracle_capi.cpp -- TEST INFRASTRUCTURE ONLY.
//
// extern "C" entry points over the CPU restatement in hg_oracle.hpp, loaded
// with ctypes by the test-suite, the smoke check and the CPU-baseline leg of
// the benchmark.  The product library never links this file.
#include <omp.h>

#include <cstring>
#include <exception>
#include <string>

#include "hg_oracle.hpp"

using orc::Consts;
using orc::Dual;
using orc::Fields;
using orc::Mesh;
using orc::Rev;

namespace {
thread_local std::string g_err;

struct FieldPtrs {  // plain-pointer form of the parameter bundle
  const double *hstill, *hstill_ghost, *zb, *zb_ghost, *S0x, *S0y, *n, *inletQ, *exitWSE;
  const double* fmag;  // nullable: FlowResistance friction magnitudes
};

template <class T>
Fields<T> make_fields(const Mesh& m, const FieldPtrs& f) {
  Fields<T> p;
  auto cp = [](const double* src, int64_t n) {
    std::vector<T> v(n);
    for (int64_t i = 0; i < n; ++i) v[i] = T(src[i]);
    return v;
  };
  p.hstill = cp(f.hstill, m.N);
  p.hstill_ghost = cp(f.hstill_ghost, m.G);
  p.zb = cp(f.zb, m.N);
  p.zb_ghost = cp(f.zb_ghost, m.G);
  p.S0x = cp(f.S0x, m.N);
  p.S0y = cp(f.S0y, m.N);
  p.n = cp(f.n, m.N);
  p.inletQ = cp(f.inletQ, m.n_inlet);
  p.exitWSE = cp(f.exitWSE, m.n_exit);
  if (f.fmag) p.fmag = cp(f.fmag, m.N);
  return p;
}

template <class F>
int guarded(F&& fn) {
  try {
    fn();
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 1;
  }
}
}  // namespace

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }

void orc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }

int orc_max_threads() { return omp_get_max_threads(); }

// One evaluation of swe_2d_rhs (semi_discretize_swe_2D.jl:18-277).
// recompute_bed != 0 mirrors the active-zb path (:114-126): zb_ghost and S0
// are rebuilt from zb before the RHS.
int orc_rhs(const Mesh* m, const Consts* c, const FieldPtrs* f, int recompute_bed, const double* Q,
            double* dQdt, double* inlet_total_A) {
  return guarded([&] {
    Fields<double> p = make_fields<double>(*m, *f);
    if (recompute_bed) orc::update_bed_data(*m, p.zb, p.zb_ghost, p.S0x, p.S0y);
    std::vector<double> q(Q, Q + 3 * m->N), r;
    std::vector<double> tA;
    orc::swe_2d_rhs(*m, *c, p, q, r, &tA);
    std::memcpy(dQdt, r.data(), sizeof(double) * 3 * m->N);
    if (inlet_total_A) std::memcpy(inlet_total_A, tA.data(), sizeof(double) * tA.size());
  });
}

// nsteps of custom_ODE_update_cells (custom_ODE_solvers.jl:5-33), Q in/out.
int orc_run(const Mesh* m, const Consts* c, const FieldPtrs* f, int recompute_bed, double* Q, double dt,
            int64_t nsteps) {
  return guarded([&] {
    Fields<double> p = make_fields<double>(*m, *f);
    if (recompute_bed) orc::update_bed_data(*m, p.zb, p.zb_ghost, p.S0x, p.S0y);
    std::vector<double> q(Q, Q + 3 * m->N);
    for (int64_t s = 0; s < nsteps; ++s) orc::custom_ODE_update_cells(*m, *c, p, q, dt);
    std::memcpy(Q, q.data(), sizeof(double) * 3 * m->N);
  });
}

// update_bed_data (process_bed_2D.jl:46-66).
int orc_bed_data(const Mesh* m, const double* zb, double* zb_ghost, double* S0x, double* S0y) {
  return guarded([&] {
    std::vector<double> z(zb, zb + m->N), zg, sx, sy;
    orc::update_bed_data(*m, z, zg, sx, sy);
    std::memcpy(zb_ghost, zg.data(), sizeof(double) * m->G);
    std::memcpy(S0x, sx.data(), sizeof(double) * m->N);
    std::memcpy(S0y, sy.data(), sizeof(double) * m->N);
  });
}

// compute_friction_terms, Manning branch (semi_discretize_swe_2D.jl:542-547),
// element-wise over arrays of length n.
int orc_friction(const Consts* c, int64_t n, const double* h, const double* qx, const double* qy,
                 const double* mn, double* fx, double* fy) {
  return guarded([&] {
    for (int64_t i = 0; i < n; ++i) orc::friction_terms(*c, h[i], qx[i], qy[i], mn[i], fx[i], fy[i]);
  });
}

// Riemann_2D_Roe on a batch of face states (swe_2D_solvers.jl:4-164).
// states: [n][12] = xiL hsL hL huL hvL zbL xiR hsR hR huR hvR zbR; normals [n][2].
int orc_roe(int64_t n, const double* states, const double* normals, double g, double hmin, double* F) {
  return guarded([&] {
    for (int64_t k = 0; k < n; ++k) {
      const double* s = states + 12 * k;
      auto r = orc::riemann_2d_roe<double>(s[0], s[1], s[2], s[3], s[4], s[5], s[6], s[7], s[8], s[9], s[10],
                                           s[11], g, normals[2 * k], normals[2 * k + 1], hmin);
      F[3 * k] = r[0];
      F[3 * k + 1] = r[1];
      F[3 * k + 2] = r[2];
    }
  });
}

double orc_cfl_dt(const Mesh* m, const Consts* c, const double* hstill, const double* Q, double cfl) {
  return orc::cfl_dt(*m, *c, hstill, Q, cfl);
}

// Forward-mode derivative of one swe_2d_rhs evaluation along a tangent on
// (Q, n, zb, inletQ); a non-NULL dzb rebuilds the bed data from the Dual zb
// (the active-zb path, semi_discretize_swe_2D.jl:114-126).
int orc_rhs_jvp(const Mesh* m, const Consts* c, const FieldPtrs* f, const double* Q, const double* dQ,
                const double* dn, const double* dzb, const double* dQin, double* out, double* dout) {
  return guarded([&] {
    using D = Dual<1>;
    Fields<D> p = make_fields<D>(*m, *f);
    if (dn)
      for (int64_t i = 0; i < m->N; ++i) p.n[i].d[0] = dn[i];
    if (dQin)
      for (int b = 0; b < m->n_inlet; ++b) p.inletQ[b].d[0] = dQin[b];
    if (dzb) {
      for (int64_t i = 0; i < m->N; ++i) p.zb[i].d[0] = dzb[i];
      orc::update_bed_data(*m, p.zb, p.zb_ghost, p.S0x, p.S0y);
    }
    std::vector<D> q(3 * m->N), r;
    for (int64_t k = 0; k < 3 * m->N; ++k) {
      q[k] = D(Q[k]);
      if (dQ) q[k].d[0] = dQ[k];
    }
    orc::swe_2d_rhs(*m, *c, p, q, r);
    for (int64_t k = 0; k < 3 * m->N; ++k) {
      out[k] = r[k].v;
      dout[k] = r[k].d[0];
    }
  });
}

// Forward-mode (Dual, one tangent direction) run of nsteps Euler steps with
// tangents on the per-cell Manning n, the bed zb and the inlet discharges.
// Outputs the final state, its tangent and the probe loss
//   L = sum_p (xi_p + hstill_p + zb_p - target_p)^2
// with its tangent.  A NULL tangent array means "not perturbed"; a non-NULL
// dzb activates the bed-data rebuild of the active-zb path.
int orc_jvp_run(const Mesh* m, const Consts* c, const FieldPtrs* f, const double* Q0, double dt, int64_t nsteps,
                const double* dn, const double* dzb, const double* dQin, int64_t n_probe, const int32_t* probe,
                const double* target, double* Qout, double* dQout, double* loss, double* dloss) {
  return guarded([&] {
    using D = Dual<1>;
    Fields<D> p = make_fields<D>(*m, *f);
    if (dn)
      for (int64_t i = 0; i < m->N; ++i) p.n[i].d[0] = dn[i];
    if (dQin)
      for (int b = 0; b < m->n_inlet; ++b) p.inletQ[b].d[0] = dQin[b];
    if (dzb) {
      for (int64_t i = 0; i < m->N; ++i) p.zb[i].d[0] = dzb[i];
      orc::update_bed_data(*m, p.zb, p.zb_ghost, p.S0x, p.S0y);
    }
    std::vector<D> q(3 * m->N);
    for (int64_t k = 0; k < 3 * m->N; ++k) q[k] = D(Q0[k]);
    for (int64_t s = 0; s < nsteps; ++s) orc::custom_ODE_update_cells(*m, *c, p, q, dt);
    for (int64_t k = 0; k < 3 * m->N; ++k) {
      Qout[k] = q[k].v;
      dQout[k] = q[k].d[0];
    }
    D L(0.0);
    for (int64_t k = 0; k < n_probe; ++k) {
      const int32_t i = probe[k];
      D r = q[i] + p.hstill[i] + p.zb[i] - target[k];
      L = L + r * r;
    }
    *loss = L.v;
    *dloss = L.d[0];
  });
}

// Forward-mode run from an initial-state tangent dQ0 (no parameter tangents):
// nsteps Euler steps of (Q0, dQ0).
int orc_jvp_state(const Mesh* m, const Consts* c, const FieldPtrs* f, const double* Q0, const double* dQ0, double dt,
                  int64_t nsteps, double* Qout, double* dQout) {
  return guarded([&] {
    using D = Dual<1>;
    Fields<D> p = make_fields<D>(*m, *f);
    std::vector<D> q(3 * m->N);
    for (int64_t k = 0; k < 3 * m->N; ++k) {
      q[k] = D(Q0[k]);
      q[k].d[0] = dQ0[k];
    }
    for (int64_t s = 0; s < nsteps; ++s) orc::custom_ODE_update_cells(*m, *c, p, q, dt);
    for (int64_t k = 0; k < 3 * m->N; ++k) {
      Qout[k] = q[k].v;
      dQout[k] = q[k].d[0];
    }
  });
}

// Frees the calling thread's tape (orc_grad_run keeps its memory for the
// next call, as a tape library keeps its tape between gradients).
void orc_release_tape() { orc::tape() = orc::Tape(); }

// Reverse-mode (tape) gradient of the probe loss
//   L = sum_p (xi_p + hstill_p + zb_p - target_p)^2
// at the final state of nsteps Euler steps from Q0, with respect to the
// per-cell Manning n, the inlet discharges, the initial state and the bed zb
// (a non-NULL gzb takes the active-zb path: zb_ghost and S0 rebuilt from zb,
// as orc_jvp_run does for a non-NULL dzb): the forward
// run keeps every state, then each step, last to first, is recorded on the
// tape from its stored input and swept backwards from the next state's
// adjoint.  NULL outputs are not formed.  Runs on one thread (the tape is per
// thread); the OpenMP thread count is restored afterwards.
int orc_grad_run(const Mesh* m, const Consts* c, const FieldPtrs* f, const double* Q0, double dt, int64_t nsteps,
                 int64_t n_probe, const int32_t* probe, const double* target, double* loss, double* gn,
                 double* gQin, double* gQ0, double* gzb) {
  const int nt = omp_get_max_threads();
  omp_set_num_threads(1);
  const int rc = guarded([&] {
    const int64_t N = m->N, M3 = 3 * N;
    if (nsteps < 0 || n_probe < 0) throw std::invalid_argument("orc_grad_run: negative nsteps or n_probe");
    for (int64_t k = 0; k < n_probe; ++k)
      if (probe[k] < 0 || probe[k] >= N) throw std::invalid_argument("orc_grad_run: probe cell out of range");
    Fields<double> pd = make_fields<double>(*m, *f);
    if (gzb) orc::update_bed_data(*m, pd.zb, pd.zb_ghost, pd.S0x, pd.S0y);
    std::vector<std::vector<double>> traj(size_t(nsteps) + 1);
    traj[0].assign(Q0, Q0 + M3);
    for (int64_t s = 0; s < nsteps; ++s) {
      traj[s + 1] = traj[s];
      orc::custom_ODE_update_cells(*m, *c, pd, traj[s + 1], dt);
    }
    std::vector<double> lam(M3, 0.0);
    double L = 0.0;
    for (int64_t k = 0; k < n_probe; ++k) {
      const int32_t i = probe[k];
      const double r = traj[nsteps][i] + pd.hstill[i] + pd.zb[i] - target[k];
      L = L + r * r;
      lam[i] += 2.0 * r;
    }
    *loss = L;
    if (gzb) {  // the loss reads zb directly as well
      std::memset(gzb, 0, sizeof(double) * N);
      for (int64_t k = 0; k < n_probe; ++k) {
        const int32_t i = probe[k];
        gzb[i] += 2.0 * (traj[nsteps][i] + pd.hstill[i] + pd.zb[i] - target[k]);
      }
    }
    if (gn) std::memset(gn, 0, sizeof(double) * N);
    if (gQin) std::memset(gQin, 0, sizeof(double) * m->n_inlet);
    Fields<Rev> p = make_fields<Rev>(*m, *f);
    orc::Tape& t = orc::tape();
    std::vector<Rev> q(M3);
    std::vector<int32_t> qin(M3);
    for (int64_t s = nsteps - 1; s >= 0; --s) {
      t.clear();
      for (int64_t k = 0; k < M3; ++k) {
        q[k] = Rev::input(traj[s][k]);
        qin[k] = q[k].i;
      }
      if (gn)
        for (int64_t i = 0; i < N; ++i) p.n[i] = Rev::input(pd.n[i]);
      if (gQin)
        for (int b = 0; b < m->n_inlet; ++b) p.inletQ[b] = Rev::input(pd.inletQ[b]);
      if (gzb) {  // the active-zb path: the bed data rebuilt from zb
        for (int64_t i = 0; i < N; ++i) p.zb[i] = Rev::input(pd.zb[i]);
        orc::update_bed_data(*m, p.zb, p.zb_ghost, p.S0x, p.S0y);
      }
      orc::custom_ODE_update_cells(*m, *c, p, q, dt);
      t.adj.assign(t.size(), 0.0);
      for (int64_t k = 0; k < M3; ++k)
        if (q[k].i >= 0) t.adj[q[k].i] += lam[k];
      t.sweep();
      for (int64_t k = 0; k < M3; ++k) lam[k] = t.adj[qin[k]];
      if (gn)
        for (int64_t i = 0; i < N; ++i) gn[i] += t.adj[p.n[i].i];
      if (gQin)
        for (int b = 0; b < m->n_inlet; ++b) gQin[b] += t.adj[p.inletQ[b].i];
      if (gzb)
        for (int64_t i = 0; i < N; ++i) gzb[i] += t.adj[p.zb[i].i];
    }
    if (gQ0) std::memcpy(gQ0, lam.data(), sizeof(double) * M3);
  });
  omp_set_num_threads(nt);
  return rc;
}

}  // extern "C"
