// pybind11 module gpumap_amd._core: the reference's _core surface
// (cpp/bindings/binding.cu:45-116, gpumap_wrapper.cuh, numpy_utils.cuh) as a
// thin layer over the C ABI in include/gpumap_c.h.  Device arrays leave
// through DLPack (kDLROCM) instead of __cuda_array_interface__
// (cai_wrapper.cuh:39-62), and each exported tensor keeps the session's
// device memory alive.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <array>
#include <cctype>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/gpumap_c.h"
#include "dlpack_abi.hpp"

namespace py = pybind11;

namespace {

[[noreturn]] void raise_status(int rc) {
  const std::string msg = gm_last_error();
  if (rc == GM_ERR_OOM) {
    PyErr_SetString(PyExc_MemoryError, msg.c_str());
    throw py::error_already_set();
  }
  if (rc == GM_ERR_INVALID && msg.find("n_features mismatch") == std::string::npos) {
    throw py::value_error(msg);
  }
  throw std::runtime_error(msg);
}

inline void check(int rc) {
  if (rc != GM_OK)
    raise_status(rc);
}

struct Params {
  gm_params p;
  Params() { gm_default_params(&p); }
};

struct Result {
  gm_result r;
};

std::string metric_to_string(int32_t m) {
  return m == GM_COSINE ? "cosine" : "euclidean";
}

int32_t metric_from_object(const py::object &v) {
  if (py::isinstance<py::str>(v)) {
    std::string s = v.cast<std::string>();
    for (char &ch : s)
      ch = (char)std::tolower((unsigned char)ch);
    if (s == "euclidean")
      return GM_EUCLIDEAN;
    if (s == "cosine")
      return GM_COSINE;
    throw py::value_error("Unknown metric '" + s +
                          "' -- expected 'euclidean' or 'cosine'");
  }
  const int m = v.cast<int>();
  if (m != GM_EUCLIDEAN && m != GM_COSINE)
    throw py::value_error("Unknown metric " + std::to_string(m));
  return m;
}

std::string knn_mode_to_string(int32_t m) {
  switch (m) {
  case GM_KNN_NNDESCENT:
    return "nndescent";
  case GM_KNN_EXACT_BF16:
    return "exact_bf16";
  case GM_KNN_EXACT_FP8:
    return "exact_fp8";
  case GM_KNN_NNDESCENT_BF16:
    return "nndescent_bf16";
  default:
    return "exact_fp32";
  }
}

int32_t knn_mode_from_object(const py::object &v) {
  if (py::isinstance<py::str>(v)) {
    std::string s = v.cast<std::string>();
    for (char &ch : s)
      ch = (char)std::tolower((unsigned char)ch);
    if (s == "nndescent" || s == "nn_descent")
      return GM_KNN_NNDESCENT;
    if (s == "exact" || s == "exact_fp32")
      return GM_KNN_EXACT_FP32;
    if (s == "exact_bf16")
      return GM_KNN_EXACT_BF16;
    if (s == "exact_fp8")
      return GM_KNN_EXACT_FP8;
    if (s == "nndescent_bf16")
      return GM_KNN_NNDESCENT_BF16;
    throw py::value_error("Unknown knn_mode '" + s +
                          "' -- expected 'nndescent', 'exact_fp32', "
                          "'exact_bf16', 'exact_fp8' or 'nndescent_bf16'");
  }
  const int m = v.cast<int>();
  if (m < GM_KNN_NNDESCENT || m > GM_KNN_NNDESCENT_BF16)
    throw py::value_error("Unknown knn_mode " + std::to_string(m));
  return m;
}

struct Session {
  gm_handle *h = nullptr;
  explicit Session(const Params &p) { check(gm_create(&p.p, &h)); }
  ~Session() {
    if (h)
      gm_destroy(h);
  }
};

void capsule_destructor(PyObject *cap) {
  // a capsule that no consumer renamed still owns its tensor
  if (cap && PyCapsule_IsValid(cap, "dltensor")) {
    auto *mt = static_cast<gm::DLManagedTensor *>(
        PyCapsule_GetPointer(cap, "dltensor"));
    if (mt && mt->deleter)
      mt->deleter(mt);
  }
}

// A device array of a session: shape and dtype plus the DLPack protocol.
struct DeviceArray {
  std::shared_ptr<Session> s;
  int32_t which;
  gm_view v;

  py::tuple shape() const {
    return py::make_tuple((py::ssize_t)v.rows, (py::ssize_t)v.cols);
  }
  std::string typestr() const {
    return std::string(v.dtype_code == 1 ? "<u" : "<f") +
           std::to_string(v.dtype_bits / 8);
  }
  py::object dlpack(py::object /*stream*/, py::kwargs) const {
    void *mt = nullptr;
    check(gm_export_dlpack(s->h, which, &mt));
    PyObject *cap = PyCapsule_New(mt, "dltensor", capsule_destructor);
    if (!cap) {
      auto *t = static_cast<gm::DLManagedTensor *>(mt);
      t->deleter(t);
      throw py::error_already_set();
    }
    return py::reinterpret_steal<py::object>(cap);
  }
  py::tuple dlpack_device() const { return py::make_tuple(10, v.device_id); }
};

struct Model {
  Params params;
  std::shared_ptr<Session> s;
  explicit Model(const Params &p) : params(p), s(std::make_shared<Session>(p)) {}

  void initialize(uint64_t n_instances, uint64_t n_features) {
    check(gm_initialize(s->h, n_instances, n_features));
  }

  void append_data(py::array x) {
    py::buffer_info info = x.request();
    if (info.ndim != 2)
      throw std::runtime_error("Expected 2D numpy array.");
    if (info.format != py::format_descriptor<float>::format())
      throw std::runtime_error("Only float32 numpy arrays are supported.");
    if (info.strides[1] != (py::ssize_t)sizeof(float) ||
        info.strides[0] != (py::ssize_t)sizeof(float) * info.shape[1])
      throw std::runtime_error(
          "Require C-contiguous numpy array (row-major float32).");
    if ((uint64_t)info.shape[1] != params.p.n_features)
      throw std::runtime_error(
          "n_features mismatch with numpy array second dimension.");
    if (info.shape[0] <= 0)
      return;
    check(gm_append_data(s->h, static_cast<const float *>(info.ptr),
                         (uint64_t)info.shape[0], (uint64_t)info.shape[1]));
  }

  // rows already in device memory (row-major float32 at `ptr`); the caller
  // has finished writing them
  void append_data_device(uintptr_t ptr, uint64_t rows, uint64_t cols) {
    if (cols != params.p.n_features)
      throw std::runtime_error(
          "n_features mismatch with numpy array second dimension.");
    check(gm_append_data_device(s->h, reinterpret_cast<const float *>(ptr),
                                rows, cols));
  }

  Result run(uint32_t ni, uint32_t nu, uint32_t ne) {
    Result r;
    int rc;
    {
      py::gil_scoped_release nogil;
      rc = gm_run(s->h, ni, nu, ne, &r.r);
    }
    check(rc);
    return r;
  }

  void close() { check(gm_close(s->h)); }

  // frees the session's device memory once no exported tensor holds it
  void release() { check(gm_release(s->h)); }
  bool released() const {
    int32_t r = 0;
    check(gm_is_released(s->h, &r));
    return r != 0;
  }

  DeviceArray get(int32_t which) const {
    DeviceArray a{s, which, {}};
    check(gm_get_view(s->h, which, &a.v));
    return a;
  }
};

// One cycle of several sessions together (gm_run_batch); ops holds one
// (insertion, update, embedding) triple per model.
std::vector<Result> run_batch(const std::vector<Model *> &models,
                              const std::vector<std::array<uint32_t, 3>> &ops) {
  if (models.size() != ops.size() || models.empty())
    throw std::invalid_argument("run_batch needs one op triple per model");
  std::vector<gm_handle *> hs;
  std::vector<uint32_t> flat;
  for (size_t i = 0; i < models.size(); ++i) {
    hs.push_back(models[i]->s->h);
    flat.insert(flat.end(), ops[i].begin(), ops[i].end());
  }
  std::vector<gm_result> rs(models.size());
  int rc;
  {
    py::gil_scoped_release nogil;
    rc = gm_run_batch(hs.data(), (int32_t)hs.size(), flat.data(), rs.data());
  }
  check(rc);
  std::vector<Result> out(models.size());
  for (size_t i = 0; i < rs.size(); ++i)
    out[i].r = rs[i];
  return out;
}

} // namespace

PYBIND11_MODULE(_core, m) {
  m.doc() = "gpumap_amd gfx950 engine (C ABI binding)";

  py::class_<Params>(m, "GPUMAPParams")
      .def(py::init<>())
      .def_property(
          "n_neighbors", [](const Params &p) { return p.p.n_neighbors; },
          [](Params &p, uint64_t v) { p.p.n_neighbors = v; })
      .def_property(
          "n_components", [](const Params &p) { return p.p.n_components; },
          [](Params &p, uint64_t v) { p.p.n_components = v; })
      .def_property(
          "min_dist", [](const Params &p) { return p.p.min_dist; },
          [](Params &p, float v) { p.p.min_dist = v; })
      .def_property(
          "negative_sample_rate",
          [](const Params &p) { return p.p.negative_sample_rate; },
          [](Params &p, uint64_t v) { p.p.negative_sample_rate = v; })
      .def_property(
          "max_epoch", [](const Params &p) { return p.p.max_epoch; },
          [](Params &p, uint64_t v) { p.p.max_epoch = v; })
      .def_property(
          "a", [](const Params &p) { return p.p.a; },
          [](Params &p, float v) { p.p.a = v; })
      .def_property(
          "b", [](const Params &p) { return p.p.b; },
          [](Params &p, float v) { p.p.b = v; })
      .def_property(
          "alpha", [](const Params &p) { return p.p.alpha; },
          [](Params &p, float v) { p.p.alpha = v; })
      .def_property(
          "verbose", [](const Params &p) { return p.p.verbose != 0; },
          [](Params &p, bool v) { p.p.verbose = v ? 1 : 0; })
      .def_property(
          "metric",
          [](const Params &p) { return metric_to_string(p.p.metric); },
          [](Params &p, const py::object &v) {
            p.p.metric = metric_from_object(v);
          })
      .def_property(
          "knn_mode",
          [](const Params &p) { return knn_mode_to_string(p.p.knn_mode); },
          [](Params &p, const py::object &v) {
            p.p.knn_mode = knn_mode_from_object(v);
          })
      .def_property(
          "n_instances", [](const Params &p) { return p.p.n_instances; },
          [](Params &p, uint64_t v) { p.p.n_instances = v; })
      .def_property(
          "n_features", [](const Params &p) { return p.p.n_features; },
          [](Params &p, uint64_t v) { p.p.n_features = v; });

  py::class_<Result>(m, "GPUMAPResult")
      .def_property_readonly("n_insertion_ops",
                             [](const Result &r) { return r.r.n_insertion_ops; })
      .def_property_readonly("n_update_ops",
                             [](const Result &r) { return r.r.n_update_ops; })
      .def_property_readonly("n_embedding_ops",
                             [](const Result &r) { return r.r.n_embedding_ops; })
      .def_property_readonly("insertion_time",
                             [](const Result &r) { return r.r.insertion_time; })
      .def_property_readonly("update_time",
                             [](const Result &r) { return r.r.update_time; })
      .def_property_readonly("embedding_time",
                             [](const Result &r) { return r.r.embedding_time; })
      .def_property_readonly(
          "insertion_completed",
          [](const Result &r) { return r.r.insertion_completed != 0; })
      .def_property_readonly(
          "update_completed",
          [](const Result &r) { return r.r.update_completed != 0; })
      .def_property_readonly(
          "embedding_completed",
          [](const Result &r) { return r.r.embedding_completed != 0; })
      .def_property_readonly(
          "n_inserted_instances",
          [](const Result &r) { return r.r.n_inserted_instances; })
      .def_property_readonly(
          "n_update_changed_edges",
          [](const Result &r) { return r.r.n_update_changed_edges; })
      .def_property_readonly(
          "embedding_pending_edges",
          [](const Result &r) { return r.r.embedding_pending_edges; })
      .def_property_readonly(
          "insertion_queue_size",
          [](const Result &r) { return r.r.insertion_queue_size; })
      .def_property_readonly(
          "update_queue_size",
          [](const Result &r) { return r.r.update_queue_size; })
      .def_property_readonly(
          "embedding_queue_size",
          [](const Result &r) { return r.r.embedding_queue_size; });
  m.attr("GPUMAPIterationResult") = m.attr("GPUMAPResult");

  py::class_<DeviceArray>(m, "DeviceArray")
      .def_property_readonly("shape", &DeviceArray::shape)
      .def_property_readonly("typestr", &DeviceArray::typestr)
      .def("__dlpack__", &DeviceArray::dlpack, py::arg("stream") = py::none())
      .def("__dlpack_device__", &DeviceArray::dlpack_device);

  py::class_<Model>(m, "GPUMAP")
      .def(py::init<const Params &>(), py::arg("config"))
      .def("initialize", &Model::initialize, py::arg("n_instances"),
           py::arg("n_features"))
      .def("append_data", &Model::append_data, py::arg("x"))
      .def("append_data_device", &Model::append_data_device, py::arg("ptr"),
           py::arg("rows"), py::arg("cols"))
      .def("run", &Model::run, py::arg("n_jobs_insertion"),
           py::arg("n_jobs_update"), py::arg("n_jobs_embedding"))
      .def("close", &Model::close)
      .def("release", &Model::release,
           "Close the session and free its device memory as soon as no "
           "exported tensor holds it; later calls on it raise RuntimeError")
      .def_property_readonly("released", &Model::released)
      .def("get_embedding", [](const Model &mm) { return mm.get(GM_EMBEDDING); })
      .def("get_neighbor_indices",
           [](const Model &mm) { return mm.get(GM_NEIGHBOR_INDICES); })
      .def("get_neighbor_dists",
           [](const Model &mm) { return mm.get(GM_NEIGHBOR_DISTS); });

  m.def("shutdown", &gm_shutdown,
        "Free every live session's device memory while the HIP runtime is "
        "up (the package's exit hook); later session calls raise "
        "RuntimeError",
        py::call_guard<py::gil_scoped_release>());
  m.def("run_batch", &run_batch, py::arg("models"), py::arg("ops"),
        "One cycle of several sessions together: each runs its op triple "
        "as GPUMAP.run would, with the same results; the device work of "
        "the sessions overlaps and their SGDs share one launch.");
}
