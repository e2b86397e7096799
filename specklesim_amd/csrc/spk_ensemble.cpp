// spk_ensemble.cpp -- a speckle ensemble on one GPU: many surface
// realisations, each solved for many incidence angles, reduced to far-field
// intensities.  The reference runs one realisation and one angle at a time
// (the angle loop of main.f90:450-463 around lib_math_solver_gmres_run with
// one precalculation per mesh); here host worker threads each own a handle
// and its HIP stream and take realisations in order, so one realisation's
// host precalculation (RWG topology, octree, lists) runs while another's
// kernels occupy the GPU, and small meshes that do not fill the GPU alone
// run side by side.  Each realisation's angles are solved in batches of up
// to 16 right-hand sides (spk_gmres_batch_device).  Built on the public
// C-ABI only.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "speckle_gpu.h"

namespace spk {
int set_error(int code, const std::string &msg);
}

namespace {

constexpr int kMaxBatch = 16;

struct Worker {
  spk_handle *h = nullptr;
  std::vector<void *> b, x;  // device vectors, 2N complex128 each
  size_t bytes = 0;
  ~Worker() {
    for (void *p : b) spk_device_free(p);
    for (void *p : x) spk_device_free(p);
    if (h) spk_destroy(h);
  }
  int vectors(int device, size_t nbytes, int R) {
    if (bytes != nbytes) {
      for (void *p : b) spk_device_free(p);
      for (void *p : x) spk_device_free(p);
      b.clear();
      x.clear();
      bytes = nbytes;
    }
    while ((int)b.size() < R) {
      void *pb = nullptr, *px = nullptr;
      int rc = spk_device_alloc(device, nbytes, &pb);
      if (rc) return rc;
      b.push_back(pb);
      rc = spk_device_alloc(device, nbytes, &px);
      if (rc) return rc;
      x.push_back(px);
    }
    return SPK_OK;
  }
};

// One realisation: precalculation, then its angles in batches.
int solve_mesh(Worker &w, int device, const spk_mesh &mesh, int n_angles, const double *k_in,
               const double *e_in, int npts, const double *r, const spk_gmres_params *gp,
               double *intensity, spk_gmres_info *info) {
  int rc = spk_setup_tri(w.h, mesh.np, mesh.pp, mesh.nt, mesh.tt);
  if (rc) return rc;
  int64_t N = 0;
  if ((rc = spk_num_edges(w.h, &N))) return rc;
  const int R = std::min(kMaxBatch, n_angles);
  if ((rc = w.vectors(device, 2 * (size_t)N * 16, R))) return rc;
  std::vector<double> E(6 * (size_t)npts), H(6 * (size_t)npts);
  for (int a0 = 0; a0 < n_angles; a0 += kMaxBatch) {
    const int nb = std::min(kMaxBatch, n_angles - a0);
    for (int i = 0; i < nb; ++i)
      if ((rc = spk_rhs_plane_wave_device(w.h, k_in + 3 * (a0 + i), e_in + 3 * (a0 + i), w.b[i])))
        return rc;
    if ((rc = spk_gmres_batch_device(w.h, nb, w.b.data(), w.x.data(), gp, info + a0))) return rc;
    for (int i = 0; i < nb; ++i) {
      if (npts == 0) continue;
      if ((rc = spk_scattered_field_device(w.h, npts, r, w.x[i], E.data(), H.data()))) return rc;
      double *out = intensity + (size_t)(a0 + i) * npts;
      for (int p = 0; p < npts; ++p) {
        double s = 0.0;
        for (int c = 0; c < 6; ++c) s += E[6 * (size_t)p + c] * E[6 * (size_t)p + c];
        out[p] = s;
      }
    }
  }
  return SPK_OK;
}

}  // namespace

extern "C" int spk_ensemble_intensity(const spk_params *params, int n_mesh, const spk_mesh *meshes,
                                      int n_angles, const double *k_in, const double *e_in,
                                      int npts, const double *r, const spk_gmres_params *gp,
                                      int n_workers, double *intensity, spk_gmres_info *info) {
  if (!params || !meshes || !k_in || !e_in || !gp || !intensity || !info || n_mesh < 0 ||
      n_angles < 0 || npts < 0 || (npts > 0 && !r))
    return spk::set_error(SPK_ERR_ARG, "bad argument");
  if (n_workers < 1 || n_workers > 16) return spk::set_error(SPK_ERR_ARG, "n_workers must be 1..16");
  for (int i = 0; i < n_mesh; ++i)
    if (!meshes[i].pp || !meshes[i].tt || meshes[i].np < 3 || meshes[i].nt < 1)
      return spk::set_error(SPK_ERR_ARG, "bad mesh " + std::to_string(i));
  if (n_mesh == 0 || n_angles == 0) return SPK_OK;
  const int nw = std::min(n_workers, n_mesh);
  std::vector<Worker> workers(nw);
  for (auto &w : workers) {
    int rc = spk_create(params, &w.h);
    if (rc) return rc;
    // a stream capture in one worker thread would reject the other workers'
    // synchronous copies (HIP ties them to the legacy stream)
    if ((rc = spk_set_inner_graph(w.h, 0))) return rc;
  }
  // tuning on realisation 0 (gp->inner_iterations < 0), applied to every
  // worker; the solves then use the tuned inner length
  spk_gmres_params g = *gp;
  if (gp->inner_iterations < 0) {
    if (gp->preconditioning != SPK_PRECOND_NEAR_FIELD)
      return spk::set_error(SPK_ERR_ARG, "tuning needs the near-field preconditioned solver");
    g.inner_iterations = 0;
    int rc = spk_setup_tri(workers[0].h, meshes[0].np, meshes[0].pp, meshes[0].nt, meshes[0].tt);
    spk_tuning tu;
    if (!rc) rc = spk_autotune(workers[0].h, std::min(kMaxBatch, n_angles), k_in, e_in, &g, &tu);
    for (auto &w : workers)
      if (!rc) rc = spk_set_tuning(w.h, tu.tiles_per_group, tu.inner_iterations);
    if (rc) return spk::set_error(rc, std::string("tuning on realisation 0: ") + spk_last_error());
  }
  gp = &g;
  std::atomic<int> next(0);
  std::atomic<bool> stop(false);
  std::mutex mu;
  int err_rc = SPK_OK;
  std::string err_msg;
  auto run = [&](int wi) {
    Worker &w = workers[wi];
    for (;;) {
      if (stop.load()) return;
      const int m = next.fetch_add(1);
      if (m >= n_mesh) return;
      const int rc = solve_mesh(w, params->device, meshes[m], n_angles, k_in, e_in, npts, r, gp,
                                intensity + (size_t)m * n_angles * npts,
                                info + (size_t)m * n_angles);
      if (rc) {
        std::lock_guard<std::mutex> lock(mu);
        if (err_rc == SPK_OK) {
          err_rc = rc;
          err_msg = "realisation " + std::to_string(m) + ": " + spk_last_error();
        }
        stop.store(true);
        return;
      }
    }
  };
  std::vector<std::thread> th;
  for (int i = 1; i < nw; ++i) th.emplace_back(run, i);
  run(0);
  for (auto &t : th) t.join();
  if (err_rc) return spk::set_error(err_rc, err_msg);
  return SPK_OK;
}
