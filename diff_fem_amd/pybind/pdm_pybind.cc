// pdm_pybind.cc — the binding a maintainer would add on the reference side:
// python/pydifem.cc's forward_all / backward_all (pydifem.cc:182-301) on top of
// libpdm's C ABI (include/pdm.h) instead of the CPU World. The reference's
// Environment would build its pdm_scene from the arrays it already holds
// (Dfobj vertices and tets, masses, attachments, muscle segments); here it is
// built from those arrays directly.
//
// g++ -O2 -std=c++17 -shared -fPIC pdm_pybind.cc -I include $(python3 -m pybind11 --includes)
//     -L diff_fem_amd -lpdm -Wl,-rpath,'$ORIGIN/..' -o pdm_pybind$(python3-config --extension-suffix)
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <stdexcept>
#include <string>
#include <vector>

#include "pdm.h"

namespace py = pybind11;

static void ck(int rc) {
  if (rc) throw std::runtime_error(pdm_last_error());
}

// One scene, one rollout, the reference's FP64 arithmetic, total_step steps
// per forward_all (Environment(total_step, config), pydifem.cc:26-36).
struct Env {
  pdm_scene s = nullptr;
  int steps = 0, n3 = 0, n_tets = 0, n_tau = 0;
  bool recorded = false;
  Env(int total_step, std::vector<double> x_rest, std::vector<int> tets, double total_mass, double h, int max_iter,
      std::vector<double> gravity, std::vector<int> attach, double attach_w, double damping) {
    if (x_rest.size() % 3 || tets.size() % 4 || gravity.size() != 3)
      throw std::invalid_argument("x_rest needs 3N values, tets 4T, gravity 3");
    pdm_scene_desc d = {};
    d.n_vertices = (int)(x_rest.size() / 3);
    d.x_rest = x_rest.data();
    d.n_tets = (int)(tets.size() / 4);
    d.tets = tets.data();
    d.mass_same = 1;
    d.total_mass = total_mass;
    d.nu_fwd = 0.4;
    d.n_attach = (int)attach.size();
    d.attach_idx = attach.data();
    d.attach_w = attach_w;
    d.h = h;
    d.max_iter = max_iter;
    d.damping = damping;
    for (int k = 0; k < 3; ++k) d.gravity[k] = gravity[k];
    d.gravity_on = 1;
    d.alpha = 1.0;
    d.mu = 1.0;
    d.max_iter1 = max_iter;
    ck(pdm_scene_create(&d, &s));
    ck(pdm_set_batch(s, 1, PDM_FP64, PDM_HOST_EARLY_EXIT, 20));
    steps = total_step;
    n3 = 3 * d.n_vertices;
    n_tets = d.n_tets;
  }
  ~Env() { pdm_scene_destroy(s); }
};

PYBIND11_MODULE(pdm_pybind, m) {
  py::class_<Env>(m, "Environment")
      .def(py::init<int, std::vector<double>, std::vector<int>, double, double, int, std::vector<double>,
                    std::vector<int>, double, double>(),
           py::arg("total_step"), py::arg("x_rest"), py::arg("tets"), py::arg("total_mass"), py::arg("h"),
           py::arg("max_iter"), py::arg("gravity"), py::arg("attach"), py::arg("attach_w"), py::arg("damping"));
  // forward_all(q, qd, tau, env) -> [q_T, qd_T]: mMu = tau[0] / (1 + tau[1]) on
  // every corotated constraint, re-factor, total_step steps (pydifem.cc:182-259)
  m.def(
      "forward_all",
      [](std::vector<double> q, std::vector<double> qd, std::vector<double> tau, Env& e) {
        if ((int)q.size() != e.n3 || (int)qd.size() != e.n3 || tau.size() < 2)
          throw std::invalid_argument("q and qd need 3N values, tau at least [E, nu]");
        std::vector<double> E(e.n_tets, tau[0]);
        ck(pdm_set_material(e.s, E.data(), tau[1]));
        std::vector<double> xT(e.n3), vT(e.n3);
        ck(pdm_forward_host(e.s, e.steps, q.data(), qd.data(), nullptr, xT.data(), vT.data(), 1));
        e.recorded = true;
        e.n_tau = (int)tau.size();
        return std::vector<std::vector<double>>{xT, vT};
      },
      py::return_value_policy::copy);
  // backward_all(q, qd, tau, dldq, dldqd, env) -> [dq0, dqd0, dtau]: the
  // gradient of <dldq, q_T> + <dldqd, qd_T> over the last forward_all, dtau[0]
  // for the one global E, dtau[1] for nu (pydifem.cc:261-301)
  m.def(
      "backward_all",
      [](std::vector<double>, std::vector<double>, std::vector<double> tau, std::vector<double> dldq,
         std::vector<double> dldqd, Env& e) {
        if (!e.recorded) throw std::runtime_error("backward_all needs a preceding forward_all");
        if ((int)dldq.size() != e.n3 || (int)dldqd.size() != e.n3)
          throw std::invalid_argument("dldq and dldqd need 3N values");
        std::vector<double> gx(e.n3), gv(e.n3), gE(e.n_tets);
        double gnu = 0;
        ck(pdm_backward_host(e.s, dldq.data(), dldqd.data(), gx.data(), gv.data(), gE.data(), &gnu, nullptr));
        std::vector<double> dtau(tau.size(), 0.0);
        for (double g : gE) dtau[0] += g;
        dtau[1] = gnu;
        e.recorded = false;
        return std::vector<std::vector<double>>{gx, gv, dtau};
      },
      py::return_value_policy::copy);
  m.attr("__version__") = "dev";
}
