// Python module boltzmann_amd._boltzmann: the reference's pyo3 API (src/lib.rs:22-411) and
// host Simulation state (src/simulation.rs:130-446, src/fields.rs:55-214) in C++, on top
// of the C ABI in include/bolt.h.  Nothing here touches HIP directly.
//
// Classes and behaviour follow the reference:
//   Omega.BGK(w) / Omega.TRT(w+, w-)                          lib.rs:22-46
//   Simulation(dev, counts, omega, q=None)                     lib.rs:187-203
//   .fluid.rho / .fluid.vel / .cells.flags / tracer.val are numpy views that alias the
//   host arrays (lib.rs:57-67), edited in Python before the first iterate and refreshed
//   after every iterate (simulation.rs:436-443)
//   iterate(n): n even; the first call sets f (and every tracer's g) to equilibrium on the
//   host in f32 and uploads everything (simulation.rs:278-286)
//   add_tracer is idempotent over the name (lib.rs:303-342); get_tracer (lib.rs:344-360)
//   set_gravity, add_boussinesq_coupling                       lib.rs:251-297
//   write_checkpoint / load_checkpoint                         lib.rs:380-398
// Errors raise boltzmann_amd.BoltError where the reference panics.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/bolt.h"
#include "checkpoint.hpp"
#include "host_init.hpp"

namespace py = pybind11;

namespace {

struct BoltError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

void check(int rc, const char* what) {
    if (rc != BOLT_OK)
        throw BoltError(std::string(what) + ": " + bolt_last_error());
}

struct OmegaV {
    int kind = BOLT_BGK;  // BOLT_BGK or BOLT_TRT
    float a = 1.f, b = 1.f;
    std::string repr() const {
        char buf[96];
        if (kind == BOLT_BGK) snprintf(buf, sizeof buf, "Omega.BGK(%g)", (double)a);
        else snprintf(buf, sizeof buf, "Omega.TRT(%g, %g)", (double)a, (double)b);
        return buf;
    }
};

int storage_from(const std::string& s) {
    if (s == "f32") return BOLT_STORE_F32;
    if (s == "f16s" || s == "f16") return BOLT_STORE_F16S;
    if (s == "bf16s" || s == "bf16") return BOLT_STORE_BF16S;
    if (s == "f64") return BOLT_STORE_F64;
    if (s == "f32f") return BOLT_STORE_F32F;
    throw BoltError("unknown storage '" + s + "' (f32, f16s, bf16s, f64, f32f)");
}

// Page-locked host storage (bolt_host_alloc) for the arrays read back after every iterate,
// so the copies run by DMA at the link's rate rather than through staged pageable copies
template <class T>
struct PinnedAlloc {
    using value_type = T;
    PinnedAlloc() = default;
    template <class U>
    PinnedAlloc(const PinnedAlloc<U>&) {}
    T* allocate(size_t n) {
        void* p = nullptr;
        if (bolt_host_alloc(n * sizeof(T), &p) != BOLT_OK || (!p && n)) throw std::bad_alloc();
        return (T*)p;
    }
    void deallocate(T* p, size_t) { (void)bolt_host_free(p); }
    template <class U>
    bool operator==(const PinnedAlloc<U>&) const { return true; }
    template <class U>
    bool operator!=(const PinnedAlloc<U>&) const { return false; }
};
using PinnedVec = std::vector<float, PinnedAlloc<float>>;

struct TracerH {
    std::string name;
    int q = 0;
    OmegaV omega;
    int dev_id = -1;
    std::vector<float> g;
    PinnedVec C;
};

struct CouplingH {
    std::string tracer;
    float alpha = 0.f, c0 = 0.f;
    std::vector<float> gravity;
};

// Host state of one simulation.  The host arrays never change size, so numpy views of
// them stay valid for the lifetime of the object.
struct Core {
    std::mutex mu;
    bolt_ctx* ctx = nullptr;
    bolt_lattice* lat = nullptr;
    std::vector<int64_t> counts;
    int d = 0, q = 0;
    OmegaV omega;
    int storage = BOLT_STORE_F32;
    std::vector<int32_t> flags;
    std::vector<float> f;
    PinnedVec rho, vel;
    std::vector<std::unique_ptr<TracerH>> tracers;
    bool has_gravity = false;
    std::vector<float> gravity;
    std::vector<CouplingH> couplings;
    uint64_t iteration = 0;
    size_t ncell = 0;

    ~Core() {
        if (lat) bolt_lattice_destroy(lat);
        if (ctx) bolt_ctx_destroy(ctx);
    }

    TracerH* find(const std::string& n) {
        for (auto& t : tracers)
            if (t->name == n) return t.get();
        return nullptr;
    }
};

// True / False / "auto" -> BOLT_ROW_ON / BOLT_ROW_OFF / BOLT_ROW_AUTO
int row_mode(const py::object& mode) {
    if (py::isinstance<py::str>(mode)) {
        if (mode.cast<std::string>() == "auto") return BOLT_ROW_AUTO;
        throw BoltError("row staging: True, False or 'auto'");
    }
    return mode.cast<bool>() ? BOLT_ROW_ON : BOLT_ROW_OFF;
}

// "stream" / "keep" / "auto" -> BOLT_CACHE_STREAM / BOLT_CACHE_KEEP / BOLT_CACHE_AUTO
int cache_policy(const std::string& policy) {
    if (policy == "stream") return BOLT_CACHE_STREAM;
    if (policy == "keep") return BOLT_CACHE_KEEP;
    if (policy == "auto") return BOLT_CACHE_AUTO;
    throw BoltError("cache policy: 'stream', 'keep' or 'auto'");
}

int device_kind(const std::string& dev) {
    std::string u;
    for (char c : dev) u.push_back((char)toupper(c));
    if (u == "GPU") return BOLT_DEVICE_GPU;
    if (u == "CPU") return BOLT_DEVICE_CPU;
    throw BoltError("Invalid device type. Expected either 'cpu' or 'gpu'.");
}

std::shared_ptr<Core> make_core(const std::string& dev, std::vector<int64_t> counts,
                                OmegaV omega, int q, int storage) {
    auto c = std::make_shared<Core>();
    c->counts = counts;
    c->d = (int)counts.size();
    c->q = q;
    c->omega = omega;
    c->storage = storage;
    size_t n = 1;
    for (auto v : counts) n *= (size_t)v;
    c->ncell = n;
    check(bolt_ctx_create(device_kind(dev), 0, &c->ctx), "device");
    check(bolt_lattice_create(c->ctx, c->d, counts.data(), q, omega.kind, omega.a,
                              omega.kind == BOLT_TRT ? omega.b : omega.a, storage, &c->lat),
          "Simulation");
    c->flags.assign(n, 0);
    c->f.assign(n * q, 0.f);
    c->rho.assign(n, 1.f);
    c->vel.assign(n * c->d, 0.f);
    return c;
}

template <typename V>
py::array view(const std::shared_ptr<Core>& core, V& v, std::vector<ssize_t> shape) {
    using T = typename V::value_type;
    std::vector<ssize_t> strides(shape.size());
    ssize_t s = sizeof(T);
    for (int i = (int)shape.size() - 1; i >= 0; --i) {
        strides[i] = s;
        s *= shape[i];
    }
    // the capsule keeps the simulation alive while the view exists
    auto* keep = new std::shared_ptr<Core>(core);
    py::capsule base(keep, [](void* p) { delete (std::shared_ptr<Core>*)p; });
    return py::array(py::dtype::of<T>(), shape, strides, v.data(), base);
}

std::vector<ssize_t> shape_of(const Core& c, int extra) {
    std::vector<ssize_t> s(c.counts.begin(), c.counts.end());
    if (extra > 0) s.push_back(extra);
    return s;
}

void write_all(Core& c) {
    check(bolt_lattice_write(c.lat, BOLT_FIELD_FLAGS, 0, BOLT_DT_I32, c.flags.data(),
                             c.flags.size() * 4), "upload flags");
    check(bolt_lattice_write(c.lat, BOLT_FIELD_RHO, 0, BOLT_DT_F32, c.rho.data(),
                             c.rho.size() * 4), "upload rho");
    check(bolt_lattice_write(c.lat, BOLT_FIELD_VEL, 0, BOLT_DT_F32, c.vel.data(),
                             c.vel.size() * 4), "upload vel");
    check(bolt_lattice_write(c.lat, BOLT_FIELD_F, 0, BOLT_DT_F32, c.f.data(), c.f.size() * 4),
          "upload f");
    for (auto& t : c.tracers) {
        check(bolt_lattice_write(c.lat, BOLT_FIELD_C, t->dev_id, BOLT_DT_F32, t->C.data(),
                                 t->C.size() * 4), "upload C");
        check(bolt_lattice_write(c.lat, BOLT_FIELD_G, t->dev_id, BOLT_DT_F32, t->g.data(),
                                 t->g.size() * 4), "upload g");
    }
}

void read_macro(Core& c, bool with_pops) {
    check(bolt_lattice_read(c.lat, BOLT_FIELD_RHO, 0, BOLT_DT_F32, c.rho.data(), c.rho.size() * 4),
          "read rho");
    check(bolt_lattice_read(c.lat, BOLT_FIELD_VEL, 0, BOLT_DT_F32, c.vel.data(), c.vel.size() * 4),
          "read vel");
    if (with_pops)
        check(bolt_lattice_read(c.lat, BOLT_FIELD_F, 0, BOLT_DT_F32, c.f.data(), c.f.size() * 4),
              "read f");
    for (auto& t : c.tracers) {
        check(bolt_lattice_read(c.lat, BOLT_FIELD_C, t->dev_id, BOLT_DT_F32, t->C.data(),
                                t->C.size() * 4), "read C");
        if (with_pops)
            check(bolt_lattice_read(c.lat, BOLT_FIELD_G, t->dev_id, BOLT_DT_F32, t->g.data(),
                                    t->g.size() * 4), "read g");
    }
}

// Simulation::iterate (simulation.rs:278-446).
void iterate(Core& c, uint64_t n, bool readback) {
    if (n % 2 != 0) throw BoltError("iters must be even");
    if (c.iteration == 0) {
        // equilibrate on the host in f32 (fields.rs:106-114, 185-193), then upload
        bolt_host::equilibrate(c.d, c.q, c.ncell, c.rho.data(), c.vel.data(), c.f.data());
        for (auto& t : c.tracers)
            bolt_host::equilibrate(c.d, t->q, c.ncell, t->C.data(), c.vel.data(), t->g.data());
        write_all(c);
        if (c.storage != BOLT_STORE_F32 && c.storage != BOLT_STORE_F32F)
            // the other formats take their initial state from the device-side equilibrium,
            // computed from the same f32 rho/vel/C in the storage's own precision
            check(bolt_lattice_equilibrate(c.lat), "equilibrate");
    }
    check(bolt_lattice_step(c.lat, n), "iterate");
    if (readback) read_macro(c, false);
    else check(bolt_lattice_sync(c.lat), "iterate");
    c.iteration += n;
}

struct PyFluid {
    std::shared_ptr<Core> core;
};
struct PyCells {
    std::shared_ptr<Core> core;
};
struct PyScalar {
    std::shared_ptr<Core> core;
    std::string name;
};
struct PySimulation {
    std::shared_ptr<Core> core;
};

struct NetCore {
    bolt_ctx* ctx = nullptr;
    bolt_nn* net = nullptr;
    int nx = 0, ny = 0, batch = 0, nfeat = 8;
    ~NetCore() {
        if (net) bolt_nn_destroy(net);
        if (ctx) bolt_ctx_destroy(ctx);
    }
};
struct PyNet {
    std::shared_ptr<NetCore> core;
};

struct BatchCore {
    bolt_ctx* ctx = nullptr;
    bolt_batch* batch = nullptr;
    std::vector<std::pair<int64_t, int64_t>> shapes;
    ~BatchCore() {
        if (batch) bolt_batch_destroy(batch);
        if (ctx) bolt_ctx_destroy(ctx);
    }
};
struct PyBatch {
    std::shared_ptr<BatchCore> core;
};

std::pair<int, int> batch_field(const std::string& f) {
    if (f == "flags") return {BOLT_FIELD_FLAGS, 1};
    if (f == "f") return {BOLT_FIELD_F, 9};
    if (f == "rho") return {BOLT_FIELD_RHO, 1};
    if (f == "vel") return {BOLT_FIELD_VEL, 2};
    if (f == "g") return {BOLT_FIELD_G, 5};
    if (f == "C") return {BOLT_FIELD_C, 1};
    throw BoltError("unknown field " + f);
}

size_t fluid_size_bytes(const Core& c) {
    return 4 * (c.f.size() + c.rho.size() + c.vel.size() +
                ((c.has_gravity || !c.couplings.empty()) ? c.ncell * c.d : 0));
}

// The Rust host keeps an acceleration buffer once gravity or a coupling exists; it only
// shows in size_bytes, as the kernels here need no per-cell acceleration array.

void add_tracer_core(Core& c, const std::string& name, OmegaV om, int q) {
    if (om.kind != BOLT_BGK)
        throw BoltError("Kernel not found: D" + std::to_string(c.d) + "Q" + std::to_string(q) +
                        " AdvectionDiffusion TRT");
    auto t = std::make_unique<TracerH>();
    t->name = name;
    t->q = q;
    t->omega = om;
    check(bolt_lattice_add_tracer(c.lat, q, om.a, &t->dev_id), "add_tracer");
    t->g.assign(c.ncell * q, 0.f);
    t->C.assign(c.ncell, 0.f);
    c.tracers.push_back(std::move(t));
}

std::vector<float> to_vec(py::array_t<float, py::array::c_style | py::array::forcecast> a) {
    auto r = a.unchecked<1>();
    std::vector<float> v(r.shape(0));
    for (ssize_t i = 0; i < r.shape(0); ++i) v[i] = r(i);
    return v;
}

}  // namespace

PYBIND11_MODULE(_boltzmann, m) {
    m.doc() = "MI355X lattice-Boltzmann update behind the reference's Simulation API";
    py::register_exception<BoltError>(m, "BoltError", PyExc_RuntimeError);

    py::class_<OmegaV>(m, "Omega")
        .def_static("BGK", [](float w) { OmegaV o; o.kind = BOLT_BGK; o.a = w; o.b = w; return o; })
        .def_static("TRT", [](float a, float b) { OmegaV o; o.kind = BOLT_TRT; o.a = a; o.b = b; return o; })
        .def("__repr__", &OmegaV::repr)
        .def_property_readonly("kind", [](const OmegaV& o) { return o.kind == BOLT_BGK ? "BGK" : "TRT"; })
        .def_property_readonly("values", [](const OmegaV& o) {
            if (o.kind == BOLT_BGK) return py::tuple(py::make_tuple(o.a));
            return py::tuple(py::make_tuple(o.a, o.b));
        });

    py::class_<PyFluid>(m, "Fluid")
        .def_property_readonly("rho", [](PyFluid& s) {
            return view(s.core, s.core->rho, shape_of(*s.core, 0));
        })
        .def_property_readonly("vel", [](PyFluid& s) {
            return view(s.core, s.core->vel, shape_of(*s.core, s.core->d));
        })
        .def_property_readonly("size_bytes", [](PyFluid& s) { return fluid_size_bytes(*s.core); });

    py::class_<PyScalar>(m, "Scalar")
        .def_property_readonly("val", [](PyScalar& s) {
            TracerH* t = s.core->find(s.name);
            if (!t) throw BoltError("No tracers configured");
            return view(s.core, t->C, shape_of(*s.core, 0));
        })
        .def_property_readonly("size_bytes", [](PyScalar& s) {
            TracerH* t = s.core->find(s.name);
            if (!t) throw BoltError("unknown tracer '" + s.name + "'");
            return 4 * (t->g.size() + t->C.size());
        })
        .def_property_readonly("name", [](PyScalar& s) { return s.name; });

    py::class_<PyCells>(m, "Cells")
        .def_property_readonly("flags", [](PyCells& s) {
            return view(s.core, s.core->flags, shape_of(*s.core, 0));
        })
        .def_property_readonly("size_bytes", [](PyCells& s) { return 4 * s.core->flags.size(); })
        .def_property_readonly("counts", [](PyCells& s) {
            py::array_t<size_t> a(s.core->counts.size());
            for (size_t i = 0; i < s.core->counts.size(); ++i) a.mutable_at(i) = (size_t)s.core->counts[i];
            return a;
        })
        .def_property_readonly("count", [](PyCells& s) { return s.core->ncell; });

    py::class_<PySimulation>(m, "Simulation")
        .def(py::init([](const std::string& dev, std::vector<int64_t> counts, OmegaV omega,
                         py::object q, const std::string& storage) {
                 const int d = (int)counts.size();
                 int qq;
                 if (q.is_none()) {
                     if (d == 1) qq = 3;
                     else if (d == 2) qq = 9;
                     else if (d == 3) qq = 27;
                     else throw BoltError("Invalid number of dimensions. Expected either 1, 2 or 3.");
                 } else {
                     qq = q.cast<int>();
                 }
                 return PySimulation{make_core(dev, counts, omega, qq, storage_from(storage))};
             }),
             py::arg("dev"), py::arg("counts"), py::arg("omega"), py::arg("q") = py::none(),
             py::kw_only(), py::arg("storage") = "f32")
        .def_property_readonly("device_info", [](PySimulation& s) {
            char buf[256];
            check(bolt_ctx_device_info(s.core->ctx, buf, sizeof buf), "device_info");
            return std::string(buf);
        })
        .def_property_readonly("size_bytes", [](PySimulation& s) {
            Core& c = *s.core;
            size_t b = fluid_size_bytes(c) + 4 * c.flags.size();
            for (auto& t : c.tracers) b += 4 * (t->g.size() + t->C.size());
            return b;
        })
        .def_property_readonly("iteration", [](PySimulation& s) { return s.core->iteration; })
        .def("iterate", [](PySimulation& s, uint64_t n) {
            std::lock_guard<std::mutex> lk(s.core->mu);
            iterate(*s.core, n, true);
        }, py::arg("iters"))
        .def("set_gravity", [](PySimulation& s, py::array_t<float, py::array::c_style | py::array::forcecast> g) {
            std::lock_guard<std::mutex> lk(s.core->mu);
            auto v = to_vec(g);
            if ((int)v.size() != s.core->d) throw BoltError("gravity must have one value per dimension");
            s.core->has_gravity = true;
            s.core->gravity = v;
            check(bolt_lattice_set_gravity(s.core->lat, v.data()), "set_gravity");
        }, py::arg("gravity"))
        .def("add_boussinesq_coupling", [](PySimulation& s, PyScalar& tracer, float alpha, float c0,
                                           py::array_t<float, py::array::c_style | py::array::forcecast> g) {
            std::lock_guard<std::mutex> lk(s.core->mu);
            TracerH* t = s.core->find(tracer.name);
            if (!t) throw BoltError("unknown tracer '" + tracer.name + "'");
            auto v = to_vec(g);
            if ((int)v.size() != s.core->d) throw BoltError("g must have one value per dimension");
            check(bolt_lattice_add_coupling(s.core->lat, t->dev_id, alpha, c0, v.data()),
                  "add_boussinesq_coupling");
            s.core->couplings.push_back({tracer.name, alpha, c0, v});
        }, py::arg("tracer"), py::arg("alpha"), py::arg("c0"), py::arg("g"))
        .def("add_tracer", [](PySimulation& s, const std::string& name, OmegaV omega, py::object q) {
            std::lock_guard<std::mutex> lk(s.core->mu);
            if (!s.core->find(name)) {
                int qq;
                if (q.is_none()) {
                    qq = s.core->d == 1 ? 3 : (s.core->d == 2 ? 5 : 7);
                } else {
                    qq = q.cast<int>();
                }
                if (s.core->iteration != 0)
                    throw BoltError("add_tracer after the first iterate is not supported");
                add_tracer_core(*s.core, name, omega, qq);
            }
            return PyScalar{s.core, name};
        }, py::arg("name"), py::arg("omega"), py::arg("q") = py::none())
        .def("get_tracer", [](PySimulation& s, const std::string& name) {
            return PyScalar{s.core, name};
        }, py::arg("name"))
        .def_property_readonly("cells", [](PySimulation& s) { return PyCells{s.core}; })
        .def_property_readonly("fluid", [](PySimulation& s) { return PyFluid{s.core}; })
        .def("write_checkpoint", [](PySimulation& s, const std::string& path) {
            std::lock_guard<std::mutex> lk(s.core->mu);
            Core& c = *s.core;
            // synchronise before serialising (the reference reads back without waiting,
            // simulation.rs:448-462)
            if (c.iteration != 0) read_macro(c, true);
            bolt_ckpt::State st;
            st.counts = c.counts;
            st.flags = &c.flags;
            st.f = &c.f;
            const std::vector<float> rho_h(c.rho.begin(), c.rho.end()), vel_h(c.vel.begin(), c.vel.end());
            st.rho = &rho_h;
            st.vel = &vel_h;
            st.q = c.q;
            st.omega_kind = c.omega.kind;
            st.omega_a = c.omega.a;
            st.omega_b = c.omega.b;
            st.has_acc = c.has_gravity || !c.couplings.empty();
            std::vector<std::vector<float>> conc;  // the tracers' C as the serialiser takes them
            conc.reserve(c.tracers.size());
            for (auto& t : c.tracers) {
                conc.emplace_back(t->C.begin(), t->C.end());
                st.tracers.push_back({t->name, t->q, t->omega.kind, t->omega.a, t->omega.b, &t->g, &conc.back()});
            }
            st.has_gravity = c.has_gravity;
            st.gravity = c.gravity;
            for (auto& cp : c.couplings) st.couplings.push_back({cp.tracer, cp.alpha, cp.c0, cp.gravity});
            st.iteration = c.iteration;
            std::string err;
            if (!bolt_ckpt::write(path, st, &err)) throw BoltError("write_checkpoint: " + err);
        }, py::arg("path"))
        .def_static("load_checkpoint", [](const std::string& dev, const std::string& path,
                                          const std::string& storage) {
            bolt_ckpt::Loaded ld;
            std::string err;
            if (!bolt_ckpt::read(path, &ld, &err)) throw BoltError("load_checkpoint: " + err);
            OmegaV om;
            om.kind = ld.omega_kind;
            om.a = ld.omega_a;
            om.b = ld.omega_b;
            auto core = make_core(dev, ld.counts, om, ld.q, storage_from(storage));
            Core& c = *core;
            c.flags = std::move(ld.flags);
            c.f = std::move(ld.f);
            c.rho.assign(ld.rho.begin(), ld.rho.end());
            c.vel.assign(ld.vel.begin(), ld.vel.end());
            for (auto& t : ld.tracers) {
                OmegaV to;
                to.kind = t.omega_kind;
                to.a = t.omega_a;
                to.b = t.omega_b;
                add_tracer_core(c, t.name, to, t.q);
                c.tracers.back()->g = std::move(t.g);
                c.tracers.back()->C.assign(t.C.begin(), t.C.end());
            }
            if (ld.has_gravity) {
                c.has_gravity = true;
                c.gravity = ld.gravity;
                check(bolt_lattice_set_gravity(c.lat, c.gravity.data()), "set_gravity");
            }
            for (auto& cp : ld.couplings) {
                TracerH* t = c.find(cp.tracer);
                if (!t) throw BoltError("load_checkpoint: coupling to unknown tracer " + cp.tracer);
                check(bolt_lattice_add_coupling(c.lat, t->dev_id, cp.alpha, cp.c0, cp.gravity.data()),
                      "add_boussinesq_coupling");
                c.couplings.push_back({cp.tracer, cp.alpha, cp.c0, cp.gravity});
            }
            c.iteration = ld.iteration;
            write_all(c);  // Simulation::load_checkpoint finalises (simulation.rs:464-479)
            return PySimulation{core};
        }, py::arg("dev"), py::arg("path"), py::kw_only(), py::arg("storage") = "f32")
        // ---- extensions beyond the reference API ----
        .def_property_readonly("storage", [](PySimulation& s) {
            const char* names[] = {"f32", "f16s", "bf16s", "f64", "f32f"};
            return std::string(names[s.core->storage]);
        })
        .def("step", [](PySimulation& s, uint64_t n) {
            // iterate without the host read-back of rho/vel/C
            std::lock_guard<std::mutex> lk(s.core->mu);
            iterate(*s.core, n, false);
        }, py::arg("iters"))
        .def("step_timed", [](PySimulation& s, uint64_t n) {
            // runs n steps (after the first-iterate initialisation has happened) and returns
            // the device time in ms from HIP events on the lattice's stream
            std::lock_guard<std::mutex> lk(s.core->mu);
            if (s.core->iteration == 0) throw BoltError("step_timed needs a prior iterate");
            double ms = 0;
            check(bolt_lattice_step_timed(s.core->lat, n, &ms), "step_timed");
            s.core->iteration += n;
            return ms;
        }, py::arg("iters"))
        .def("read_field", [](PySimulation& s, const std::string& field, py::object tracer,
                              const std::string& dtype) {
            // fresh copy of a device field; dtype "f32" or "f64"
            std::lock_guard<std::mutex> lk(s.core->mu);
            Core& c = *s.core;
            int fid, per = 1, tid = 0;
            if (field == "f") { fid = BOLT_FIELD_F; per = c.q; }
            else if (field == "rho") fid = BOLT_FIELD_RHO;
            else if (field == "vel") { fid = BOLT_FIELD_VEL; per = c.d; }
            else if (field == "g" || field == "C") {
                TracerH* t = c.find(tracer.cast<std::string>());
                if (!t) throw BoltError("unknown tracer");
                tid = t->dev_id;
                fid = field == "g" ? BOLT_FIELD_G : BOLT_FIELD_C;
                if (field == "g") per = t->q;
            } else if (field == "flags") {
                std::vector<ssize_t> shp(c.counts.begin(), c.counts.end());
                py::array_t<int32_t> out(shp);
                check(bolt_lattice_read(c.lat, BOLT_FIELD_FLAGS, 0, BOLT_DT_I32, out.mutable_data(), c.ncell * 4), "read_field");
                return py::array(out);
            } else throw BoltError("unknown field " + field);
            std::vector<ssize_t> shp(c.counts.begin(), c.counts.end());
            if (per > 1 || field == "vel") shp.push_back(per);
            if (dtype == "f64") {
                py::array_t<double> out(shp);
                check(bolt_lattice_read(c.lat, fid, tid, BOLT_DT_F64, out.mutable_data(), c.ncell * per * 8), "read_field");
                return py::array(out);
            }
            py::array_t<float> out(shp);
            check(bolt_lattice_read(c.lat, fid, tid, BOLT_DT_F32, out.mutable_data(), c.ncell * per * 4), "read_field");
            return py::array(out);
        }, py::arg("field"), py::arg("tracer") = py::none(), py::arg("dtype") = "f32")
        .def("reduce", [](PySimulation& s, const std::string& what, py::object tracer) {
            std::lock_guard<std::mutex> lk(s.core->mu);
            Core& c = *s.core;
            double out[3] = {0, 0, 0};
            if (what == "mass") {
                check(bolt_lattice_reduce(c.lat, BOLT_REDUCE_MASS, 0, out), "reduce");
                return py::cast(out[0]);
            }
            if (what == "momentum") {
                check(bolt_lattice_reduce(c.lat, BOLT_REDUCE_MOMENTUM, 0, out), "reduce");
                py::list l;
                for (int a = 0; a < c.d; ++a) l.append(out[a]);
                return py::object(l);
            }
            if (what == "tracer") {
                TracerH* t = c.find(tracer.cast<std::string>());
                if (!t) throw BoltError("unknown tracer");
                check(bolt_lattice_reduce(c.lat, BOLT_REDUCE_TRACER, t->dev_id, out), "reduce");
                return py::cast(out[0]);
            }
            throw BoltError("unknown reduction " + what);
        }, py::arg("what"), py::arg("tracer") = py::none())
        .def("residual", [](PySimulation& s, uint64_t steps_between) {
            // max|u - u_prev| / (steps_between * max|u|), FP64 on the device; the first
            // call stores the snapshot and returns inf
            std::lock_guard<std::mutex> lk(s.core->mu);
            double r = 0;
            check(bolt_lattice_residual(s.core->lat, steps_between, &r), "residual");
            return r;
        }, py::arg("steps_between"))
        .def("count_nonfinite", [](PySimulation& s, const std::string& field) {
            // non-finite values of the device rho or vel after the last iterate, counted on
            // the device (the host check of simulation.py:241-242 without its pass over the
            // host array)
            std::lock_guard<std::mutex> lk(s.core->mu);
            if (s.core->iteration == 0) throw BoltError("count_nonfinite reads the device fields: call it after iterate");
            int fid;
            if (field == "vel") fid = BOLT_FIELD_VEL;
            else if (field == "rho") fid = BOLT_FIELD_RHO;
            else throw BoltError("count_nonfinite: field must be 'vel' or 'rho'");
            uint64_t n = 0;
            check(bolt_lattice_count_nonfinite(s.core->lat, fid, &n), "count_nonfinite");
            return n;
        }, py::arg("field") = "vel")
        .def("postproc", [](PySimulation& s, const std::string& what, py::object origin) {
            // utils/postproc.py:12-127 on the device velocity (lattice units)
            std::lock_guard<std::mutex> lk(s.core->mu);
            Core& c = *s.core;
            if (c.iteration == 0)
                throw BoltError("postproc reads the device velocity: call it after iterate");
            int kind;
            if (what == "vmag") kind = BOLT_POST_VMAG;
            else if (what == "curl") kind = BOLT_POST_CURL;
            else if (what == "stream") kind = BOLT_POST_STREAM;
            else throw BoltError("unknown post-processing '" + what + "' (vmag, curl, stream)");
            std::vector<ssize_t> shape(c.counts.begin(), c.counts.end());
            if (kind == BOLT_POST_CURL && c.d == 3) shape.push_back(3);
            py::array_t<float> out(shape);
            int o[2] = {0, 0};
            if (!origin.is_none()) {
                auto t = origin.cast<std::vector<int>>();
                if (t.size() != 2) throw BoltError("origin must be (x, y)");
                o[0] = t[0];
                o[1] = t[1];
            }
            check(bolt_lattice_postproc(c.lat, kind, o, out.mutable_data(), (size_t)out.size() * 4),
                  "postproc");
            return out;
        }, py::arg("what"), py::arg("origin") = py::none())
        .def("set_block", [](PySimulation& s, int b) { check(bolt_lattice_set_block(s.core->lat, b), "set_block"); })
        .def("set_vector_width", [](PySimulation& s, int vw) {
            // cells per lane of the fluid step (0: chosen per lattice); results do not change
            check(bolt_lattice_set_vector_width(s.core->lat, vw), "set_vector_width");
        }, py::arg("vw"))
        .def("set_row_staging", [](PySimulation& s, py::object mode) {
            // the even step's kernel: True (row-staged wherever it can run), False (direct)
            // or "auto" (the default: where it measured faster); results do not change
            check(bolt_lattice_set_row_staging(s.core->lat, row_mode(mode)), "set_row_staging");
        }, py::arg("mode"))
        .def("set_cache_policy", [](PySimulation& s, const std::string& policy) {
            // population loads: "stream" (streaming), "keep" (ordinary: the lines stay in the
            // Infinity Cache) or "auto" (the default: by the step's working set); results do
            // not change
            check(bolt_lattice_set_cache_policy(s.core->lat, cache_policy(policy)), "set_cache_policy");
        }, py::arg("policy"))
        .def_property_readonly("device_bytes", [](PySimulation& s) {
            uint64_t b = 0;
            check(bolt_lattice_device_bytes(s.core->lat, &b), "device_bytes");
            return b;
        });

    // ---- batches of small coupled lattices (no counterpart in the reference API) ----
    py::class_<PyBatch>(m, "Batch")
        .def(py::init([](const std::string& dev, std::vector<std::vector<int64_t>> counts,
                         const std::string& collision, std::vector<std::vector<float>> omega_f,
                         py::object omega_g_o, py::object alpha_o, py::object c0_o,
                         py::object g_o) {
                 const size_t n = counts.size();
                 const bool tracer = !omega_g_o.is_none();
                 std::vector<float> omega_g, alpha, c0;
                 std::vector<std::vector<float>> g(n, std::vector<float>{0.f, 0.f});
                 if (tracer) {
                     omega_g = omega_g_o.cast<std::vector<float>>();
                     alpha = alpha_o.cast<std::vector<float>>();
                     c0 = c0_o.cast<std::vector<float>>();
                     g = g_o.cast<std::vector<std::vector<float>>>();
                 }
                 if (omega_f.size() != n ||
                     (tracer && (omega_g.size() != n || alpha.size() != n || c0.size() != n ||
                                 g.size() != n)))
                     throw BoltError("Batch: every per-instance list needs one entry per instance");
                 std::vector<int64_t> cc;
                 std::vector<float> of, gg;
                 std::vector<std::pair<int64_t, int64_t>> shapes;
                 for (size_t k = 0; k < n; ++k) {
                     if (counts[k].size() != 2 || g[k].size() != 2)
                         throw BoltError("Batch: instances are 2D (counts and g of length 2)");
                     cc.push_back(counts[k][0]);
                     cc.push_back(counts[k][1]);
                     shapes.push_back({counts[k][0], counts[k][1]});
                     of.push_back(omega_f[k].at(0));
                     of.push_back(omega_f[k].size() > 1 ? omega_f[k][1] : omega_f[k][0]);
                     gg.push_back(g[k][0]);
                     gg.push_back(g[k][1]);
                 }
                 int coll;
                 if (collision == "bgk") coll = BOLT_BGK;
                 else if (collision == "trt") coll = BOLT_TRT;
                 else throw BoltError("collision must be 'bgk' or 'trt'");
                 auto b = std::make_shared<BatchCore>();
                 check(bolt_ctx_create(device_kind(dev), 0, &b->ctx), "device");
                 check(bolt_batch_create(b->ctx, (int)n, cc.data(), coll, of.data(),
                                         tracer ? omega_g.data() : nullptr, tracer ? alpha.data() : nullptr,
                                         tracer ? c0.data() : nullptr, tracer ? gg.data() : nullptr,
                                         &b->batch), "Batch");
                 b->shapes = shapes;
                 return PyBatch{b};
             }),
             py::arg("dev"), py::arg("counts"), py::arg("collision"), py::arg("omega_f"),
             py::arg("omega_g") = py::none(), py::arg("alpha") = py::none(),
             py::arg("c0") = py::none(), py::arg("g") = py::none())
        .def("residual", [](PyBatch& b, int k, uint64_t steps_between) {
            double r = 0;
            check(bolt_batch_residual(b.core->batch, k, steps_between, &r), "residual");
            return r;
        }, py::arg("inst"), py::arg("steps_between"))
        .def("residuals", [](PyBatch& b, uint64_t steps_between) {
            // every active instance's residual in one launch (NaN: inactive)
            py::array_t<double> out((ssize_t)b.core->shapes.size());
            check(bolt_batch_residuals(b.core->batch, steps_between, out.mutable_data()), "residuals");
            return out;
        }, py::arg("steps_between"))
        .def_property_readonly("n", [](PyBatch& b) { return b.core->shapes.size(); })
        .def_property_readonly("cells", [](PyBatch& b) {
            uint64_t c = 0;
            check(bolt_batch_cells(b.core->batch, &c), "cells");
            return c;
        })
        .def("write", [](PyBatch& b, int k, const std::string& field, py::array arr) {
            auto [fid, per] = batch_field(field);
            const auto& sh = b.core->shapes.at(k);
            const size_t n = (size_t)sh.first * sh.second * per;
            if (fid == BOLT_FIELD_FLAGS) {
                auto a = py::array_t<int32_t, py::array::c_style | py::array::forcecast>(arr);
                if ((size_t)a.size() != n) throw BoltError("write: wrong size");
                check(bolt_batch_write(b.core->batch, k, fid, BOLT_DT_I32, a.data(), n * 4), "write");
            } else {
                auto a = py::array_t<float, py::array::c_style | py::array::forcecast>(arr);
                if ((size_t)a.size() != n) throw BoltError("write: wrong size");
                check(bolt_batch_write(b.core->batch, k, fid, BOLT_DT_F32, a.data(), n * 4), "write");
            }
        }, py::arg("inst"), py::arg("field"), py::arg("array"))
        .def("read", [](PyBatch& b, int k, const std::string& field) {
            auto [fid, per] = batch_field(field);
            const auto& sh = b.core->shapes.at(k);
            std::vector<ssize_t> shp = {(ssize_t)sh.first, (ssize_t)sh.second};
            if (per > 1 || field == "vel") shp.push_back(per);
            const size_t n = (size_t)sh.first * sh.second * per;
            if (fid == BOLT_FIELD_FLAGS) {
                py::array_t<int32_t> out(shp);
                check(bolt_batch_read(b.core->batch, k, fid, BOLT_DT_I32, out.mutable_data(), n * 4), "read");
                return py::array(out);
            }
            py::array_t<float> out(shp);
            check(bolt_batch_read(b.core->batch, k, fid, BOLT_DT_F32, out.mutable_data(), n * 4), "read");
            return py::array(out);
        }, py::arg("inst"), py::arg("field"))
        .def("equilibrate", [](PyBatch& b) { check(bolt_batch_equilibrate(b.core->batch), "equilibrate"); })
        .def("step", [](PyBatch& b, uint64_t n) { check(bolt_batch_step(b.core->batch, n), "step"); }, py::arg("iters"))
        .def("step_timed", [](PyBatch& b, uint64_t n) {
            double ms = 0;
            check(bolt_batch_step_timed(b.core->batch, n, &ms), "step_timed");
            return ms;
        }, py::arg("iters"))
        .def("set_shape", [](PyBatch& b, int block, int graph_steps) {
            check(bolt_batch_set_shape(b.core->batch, block, graph_steps), "set_shape");
        }, py::arg("block"), py::arg("graph_steps"))
        .def("set_row_staging", [](PyBatch& b, py::object mode) {
            check(bolt_batch_set_row_staging(b.core->batch, row_mode(mode)), "set_row_staging");
        }, py::arg("mode"))
        .def("set_cache_policy", [](PyBatch& b, const std::string& policy) {
            check(bolt_batch_set_cache_policy(b.core->batch, cache_policy(policy)), "set_cache_policy");
        }, py::arg("policy"))
        .def("set_active", [](PyBatch& b, py::array_t<bool, py::array::c_style | py::array::forcecast> a) {
            std::vector<uint8_t> v(a.data(), a.data() + a.size());
            check(bolt_batch_set_active(b.core->batch, v.data(), (int)v.size()), "set_active");
        }, py::arg("active"))
        .def("tune", [](PyBatch& b, int steps) {
            char rep[8192];  // one line per candidate (36) and the choice
            check(bolt_batch_tune(b.core->batch, steps, rep, sizeof rep), "tune");
            return std::string(rep);
        }, py::arg("steps") = 40)
        .def("reduce", [](PyBatch& b, int k, const std::string& what) {
            double out[3] = {0, 0, 0};
            int w = what == "mass" ? BOLT_REDUCE_MASS : what == "tracer" ? BOLT_REDUCE_TRACER : -1;
            if (w < 0) throw BoltError("reduce: 'mass' or 'tracer'");
            check(bolt_batch_reduce(b.core->batch, k, w, out), "reduce");
            return out[0];
        }, py::arg("inst"), py::arg("what"));

    // ---- the warm-start network ----
    py::class_<PyNet>(m, "Network")
        .def(py::init([](const std::string& dev, int nx, int ny, int batch, int width, uint64_t seed,
                         int depth, int features) {
                 auto n = std::make_shared<NetCore>();
                 check(bolt_ctx_create(device_kind(dev), 0, &n->ctx), "device");
                 check(bolt_nn_create(n->ctx, nx, ny, batch, width, depth, seed, &n->net), "Network");
                 check(bolt_nn_set_features(n->net, features), "Network");
                 n->nx = nx;
                 n->ny = ny;
                 n->batch = batch;
                 n->nfeat = features;
                 return PyNet{n};
             }),
             py::arg("dev"), py::arg("nx"), py::arg("ny"), py::arg("batch"), py::arg("width") = 32,
             py::arg("seed") = 0, py::arg("depth") = 3, py::arg("features") = 8)
        .def_property_readonly("features", [](PyNet& n) { return n.core->nfeat; })
        .def_property_readonly("param_count", [](PyNet& n) {
            uint64_t c = 0;
            check(bolt_nn_param_count(n.core->net, &c), "param_count");
            return c;
        })
        .def("get_params", [](PyNet& n) {
            uint64_t c = 0;
            check(bolt_nn_param_count(n.core->net, &c), "param_count");
            py::array_t<float> out((ssize_t)c);
            check(bolt_nn_get_params(n.core->net, out.mutable_data(), c), "get_params");
            return out;
        })
        .def("set_params", [](PyNet& n, py::array_t<float, py::array::c_style | py::array::forcecast> a) {
            check(bolt_nn_set_params(n.core->net, a.data(), (uint64_t)a.size()), "set_params");
        })
        .def("loss_grad", [](PyNet& n, py::array_t<float, py::array::c_style | py::array::forcecast> x,
                             py::array_t<float, py::array::c_style | py::array::forcecast> y) {
            NetCore& c = *n.core;
            const size_t pix = (size_t)c.batch * c.nx * c.ny;
            if ((size_t)x.size() != pix * c.nfeat || (size_t)y.size() != pix * 3)
                throw BoltError("loss_grad: x must be [batch, nx, ny, features] and y [batch, nx, ny, 3]");
            uint64_t np_ = 0;
            check(bolt_nn_param_count(c.net, &np_), "param_count");
            py::array_t<float> g((ssize_t)np_);
            double loss = 0;
            check(bolt_nn_loss_grad(c.net, x.data(), y.data(), &loss, g.mutable_data(), np_), "loss_grad");
            return py::make_tuple(loss, g);
        })
        .def("set_data", [](PyNet& n, py::array_t<float, py::array::c_style | py::array::forcecast> x,
                            py::array_t<float, py::array::c_style | py::array::forcecast> y) {
            NetCore& c = *n.core;
            const size_t pix = (size_t)c.nx * c.ny;
            const size_t cnt = (size_t)x.size() / (pix * c.nfeat);
            if (cnt * pix * c.nfeat != (size_t)x.size() || (size_t)y.size() != cnt * pix * 3)
                throw BoltError("set_data: x must be [n, nx, ny, features] and y [n, nx, ny, 3]");
            check(bolt_nn_set_data(c.net, (int)cnt, x.data(), y.data()), "set_data");
        })
        .def("train", [](PyNet& n, int epochs, float lr, uint64_t seed, int first, int total) {
            py::array_t<float> losses(std::max(epochs, 0));
            if (total <= 0) total = first + epochs;
            check(bolt_nn_train_part(n.core->net, first, epochs, total, lr, seed, losses.mutable_data()),
                  "train");
            return losses;
        }, py::arg("epochs"), py::arg("lr") = 1e-3f, py::arg("seed") = 0, py::arg("first") = 0,
           py::arg("total") = 0)
        .def("predict", [](PyNet& n, py::array_t<float, py::array::c_style | py::array::forcecast> x) {
            NetCore& c = *n.core;
            const size_t pix = (size_t)c.nx * c.ny;
            const size_t cnt = (size_t)x.size() / (pix * c.nfeat);
            if (cnt * pix * c.nfeat != (size_t)x.size() || cnt == 0)
                throw BoltError("predict: x must be [n, nx, ny, features]");
            py::array_t<float> y(std::vector<ssize_t>{(ssize_t)cnt, c.nx, c.ny, 3});
            check(bolt_nn_predict(c.net, (int)cnt, x.data(), y.mutable_data()), "predict");
            return y;
        });

    m.def("gemm_check", [](const std::string& dev, int a_form, int b_form,
                           py::array_t<float, py::array::c_style | py::array::forcecast> a,
                           py::array_t<float, py::array::c_style | py::array::forcecast> b, int k_split) {
        // C[M][N] = A.B^T through the network GEMM; a is [M][K] (a_form 0) or [K][M] (1),
        // b is [N][K] (b_form 0) or [K][N] (1)
        if (a.ndim() != 2 || b.ndim() != 2) throw BoltError("gemm_check: 2D arrays");
        const int M = (int)(a_form == 0 ? a.shape(0) : a.shape(1));
        const int K = (int)(a_form == 0 ? a.shape(1) : a.shape(0));
        const int N = (int)(b_form == 0 ? b.shape(0) : b.shape(1));
        const int Kb = (int)(b_form == 0 ? b.shape(1) : b.shape(0));
        if (K != Kb) throw BoltError("gemm_check: K mismatch");
        bolt_ctx* ctx = nullptr;
        check(bolt_ctx_create(device_kind(dev), 0, &ctx), "device");
        py::array_t<float> c(std::vector<ssize_t>{M, N});
        const int rc = bolt_nn_gemm_check(ctx, a_form, b_form, M, N, K, a.data(), b.data(), c.mutable_data(), k_split);
        bolt_ctx_destroy(ctx);
        check(rc, "gemm_check");
        return c;
    }, py::arg("dev"), py::arg("a_form"), py::arg("b_form"), py::arg("a"), py::arg("b"), py::arg("k_split") = 0);

    m.def("version", []() { return std::string(bolt_version()); });
}
