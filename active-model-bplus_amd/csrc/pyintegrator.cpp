// pyintegrator.cpp -- pybind11 module `activemodelbplus.integrator.integrator`.
//
// Same Python names and argument meaning as the reference's bindings
// (src/integrator.cc:31-124): Model, Stencil, Integrator and the exception
// CudaException.  Everything below calls the C ABI of libamb.so (include/amb.h);
// no HIP or torch type appears here.  Additions are keyword-only with defaults,
// so every positional call of the reference behaves the same.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <cstdint>
#include <cstring>
#include <random>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "amb.h"

namespace py = pybind11;

namespace {

// Reference: kernel::CudaError (src/integrator.h:10-14), registered as CudaException.
class DeviceError : public std::runtime_error {
   public:
    using std::runtime_error::runtime_error;
};

void check(int rc)
{
    if (rc == AMB_OK) return;
    const std::string msg = amb_last_error();
    if (rc == AMB_E_ARG) throw py::value_error(msg);
    throw DeviceError(msg);
}

struct Model {
    double a = 0, b = 0, c = 0, kappa = 0, lamb = 0, zeta = 0, T = 0;
    amb_model c_model() const { return amb_model{a, b, c, kappa, lamb, zeta, T}; }
    py::tuple as_tuple() const { return py::make_tuple(a, b, c, kappa, lamb, zeta, T); }
    bool operator==(const Model& o) const
    {
        return a == o.a && b == o.b && c == o.c && kappa == o.kappa && lamb == o.lamb && zeta == o.zeta && T == o.T;
    }
};

struct Stencil {
    double dt = 0, dx = 0, dy = 0;
    amb_stencil c_stencil() const { return amb_stencil{dt, dx, dy}; }
    py::tuple as_tuple() const { return py::make_tuple(dt, dx, dy); }
    bool operator==(const Stencil& o) const { return dt == o.dt && dx == o.dx && dy == o.dy; }
};

// Keys as the reference's model_as_dict (src/integrator.cc:18-29): note 'temperature'.
py::dict model_dict(const Model& m)
{
    py::dict d;
    d["a"] = m.a;
    d["b"] = m.b;
    d["c"] = m.c;
    d["kappa"] = m.kappa;
    d["lamb"] = m.lamb;
    d["zeta"] = m.zeta;
    d["temperature"] = m.T;
    return d;
}

py::dict stencil_dict(const Stencil& s)
{
    py::dict d;
    d["dt"] = s.dt;
    d["dx"] = s.dx;
    d["dy"] = s.dy;
    return d;
}

using Field = py::array_t<double, py::array::c_style | py::array::forcecast>;

uint64_t random_seed()
{
    std::random_device rd;
    return ((uint64_t)rd() << 32) ^ (uint64_t)rd();
}

class Integrator {
   public:
    Integrator(const Field& field, const Stencil& stencil, const Model& model, py::object seed, int precision,
               int64_t timestep, int tile)
        : stencil_(stencil), model_(model)
    {
        if (field.ndim() != 2) throw py::value_error("field must be a 2d array");
        ny_ = field.shape(0);
        nx_ = field.shape(1);
        amb_options opt{};
        opt.precision = precision;
        opt.seed = seed.is_none() ? random_seed() : seed.cast<uint64_t>();
        opt.timestep = timestep;
        opt.tile = tile;
        const amb_stencil st = stencil.c_stencil();
        const amb_model md = model.c_model();
        check(amb_create(field.data(), ny_, nx_, &st, &md, &opt, &h_));
    }
    ~Integrator() { amb_destroy(h_); }
    Integrator(const Integrator&) = delete;
    Integrator& operator=(const Integrator&) = delete;

    Stencil stencil() const { return stencil_; }
    Model model() const { return model_; }

    py::array_t<double> field() const
    {
        py::array_t<double> out({ny_, nx_});
        check(amb_get_field(h_, out.mutable_data()));
        return out;
    }

    py::array_t<double> scalar(int which) const
    {
        py::array_t<double> out({ny_, nx_});
        check(amb_get_observable(h_, which, out.mutable_data()));
        return out;
    }

    // Currents as [y, x] lists of (ny, nx) arrays, as the reference's HostCurrent.
    py::list vector(int which) const
    {
        py::array_t<double> both({(py::ssize_t)2, ny_, nx_});
        check(amb_get_observable(h_, which, both.mutable_data()));
        py::list out;
        for (py::ssize_t c = 0; c < 2; ++c) {
            py::array_t<double> comp({ny_, nx_});
            std::memcpy(comp.mutable_data(), both.data() + c * ny_ * nx_, sizeof(double) * ny_ * nx_);
            out.append(comp);
        }
        return out;
    }

    int64_t timestep() const
    {
        int64_t t = 0;
        check(amb_get_timestep(h_, &t));
        return t;
    }

    double time() const { return (double)timestep() * stencil_.dt; }

    void run(int64_t nsteps)
    {
        int rc;
        {
            py::gil_scoped_release release;
            rc = amb_run(h_, nsteps);
        }
        check(rc);
    }

    double run_timed(int64_t nsteps)
    {
        int rc;
        double ms = 0;
        {
            py::gil_scoped_release release;
            rc = amb_run_timed(h_, nsteps, &ms);
        }
        check(rc);
        return ms;
    }

    py::tuple structure_factor_bins(py::array_t<double, py::array::c_style | py::array::forcecast> edges)
    {
        if (edges.ndim() != 1 || edges.shape(0) < 1) throw py::value_error("edges must be a non-empty 1-D array");
        const int ne = (int)edges.shape(0);
        py::array_t<double> sums(ne + 1);
        py::array_t<int64_t> counts(ne + 1);
        int rc;
        {
            py::gil_scoped_release release;
            rc = amb_structure_factor(h_, edges.data(), ne, sums.mutable_data(), counts.mutable_data());
        }
        check(rc);
        return py::make_tuple(sums, counts);
    }

    py::tuple mass() const
    {
        double s = 0;
        int64_t bad = 0;
        check(amb_mass(h_, &s, &bad));
        return py::make_tuple(s, bad);
    }

    py::tuple last_mass() const
    {
        double s = 0;
        int64_t bad = 0;
        check(amb_last_mass(h_, &s, &bad));
        return py::make_tuple(s, bad);
    }

    void set_noise(const py::array_t<double, py::array::c_style | py::array::forcecast>& noise)
    {
        const int64_t per = 2 * (int64_t)ny_ * (int64_t)nx_;
        if (noise.size() % per != 0) throw py::value_error("noise must hold whole steps of 2*ny*nx normals");
        check(amb_set_noise(h_, noise.data(), (int64_t)noise.size() / per));
    }

    py::dict options() const
    {
        amb_options o{};
        check(amb_get_options(h_, &o));
        py::dict d;
        d["precision"] = o.precision;
        d["seed"] = o.seed;
        d["timestep"] = o.timestep;
        d["tile"] = o.tile;
        return d;
    }

    void set_tile(int index) { check(amb_set_tile(h_, index)); }

    void set_field(const Field& field)
    {
        if (field.ndim() != 2 || field.shape(0) != ny_ || field.shape(1) != nx_)
            throw py::value_error("field must have the integrator's shape");
        check(amb_set_field(h_, field.data()));
    }

    amb_handle* handle() const { return h_; }

    int autotune(int reps, bool cached)
    {
        int best = 0, rc;
        {
            py::gil_scoped_release release;
            rc = amb_autotune(h_, cached ? -reps : reps, &best);
        }
        check(rc);
        return best;
    }

   private:
    Stencil stencil_;
    Model model_;
    py::ssize_t ny_ = 0, nx_ = 0;
    amb_handle* h_ = nullptr;
};

using F32 = py::array_t<float, py::array::c_style | py::array::forcecast>;

// Learned flux-form surrogate over amb_surrogate_* (no reference counterpart).
class Surrogate {
   public:
    Surrogate(int64_t ny, int64_t nx, int max_batch, uint64_t seed) : ny_(ny), nx_(nx), max_batch_(max_batch)
    {
        check(amb_surrogate_create(ny, nx, max_batch, seed, &s_));
        check(amb_surrogate_param_count(s_, &np_));
    }
    ~Surrogate() { amb_surrogate_destroy(s_); }
    Surrogate(const Surrogate&) = delete;
    Surrogate& operator=(const Surrogate&) = delete;

    int64_t param_count() const { return np_; }

    py::array_t<float> get_params() const
    {
        py::array_t<float> out(np_);
        check(amb_surrogate_get_params(s_, out.mutable_data()));
        return out;
    }

    void set_params(const F32& p)
    {
        if (p.size() != np_) throw py::value_error("wrong parameter count");
        check(amb_surrogate_set_params(s_, p.data()));
    }

    int batch_of(const Field& phi, const Field& params) const
    {
        if (phi.ndim() != 3 || phi.shape(1) != ny_ || phi.shape(2) != nx_)
            throw py::value_error("fields must have shape (batch, ny, nx)");
        if (params.ndim() != 2 || params.shape(0) != phi.shape(0) || params.shape(1) != 3)
            throw py::value_error("params must have shape (batch, 3): lamb, zeta, phi0");
        return (int)phi.shape(0);
    }

    py::tuple loss_grad(const Field& phi, const Field& params, const Field& target)
    {
        const int b = batch_of(phi, params);
        if (target.size() != phi.size()) throw py::value_error("target must match phi");
        double loss = 0;
        py::array_t<float> grad(np_);
        check(amb_surrogate_loss_grad(s_, phi.data(), params.data(), target.data(), b, &loss, grad.mutable_data()));
        return py::make_tuple(loss, grad);
    }

    void set_dataset(const F32& phi_in, const F32& phi_out, const F32& params)
    {
        if (phi_in.ndim() != 3 || phi_in.shape(1) != ny_ || phi_in.shape(2) != nx_ || phi_out.size() != phi_in.size() ||
            params.size() != 3 * phi_in.shape(0))
            throw py::value_error("dataset must be (count, ny, nx) x2 and (count, 3)");
        check(amb_surrogate_set_dataset(s_, phi_in.data(), phi_out.data(), params.data(), phi_in.shape(0)));
    }

    void set_train_noise(double rel, uint64_t seed) { check(amb_surrogate_set_train_noise(s_, rel, seed)); }

    double train(int64_t iterations, int batch, double lr, double clip)
    {
        double loss = 0;
        int rc;
        {
            py::gil_scoped_release release;
            rc = amb_surrogate_train(s_, iterations, batch, lr, clip, &loss);
        }
        check(rc);
        return loss;
    }

    py::tuple predict(const Field& phi, const Field& params)
    {
        const int b = batch_of(phi, params);
        py::array_t<double> nxt({(py::ssize_t)b, (py::ssize_t)ny_, (py::ssize_t)nx_});
        py::array_t<double> cur({(py::ssize_t)b, (py::ssize_t)ny_, (py::ssize_t)nx_, (py::ssize_t)2});
        check(amb_surrogate_predict(s_, phi.data(), params.data(), b, nxt.mutable_data(), cur.mutable_data()));
        return py::make_tuple(nxt, cur);
    }

    double advance(py::object integrator, int64_t nsteps, const std::vector<double>& params3, int64_t euler_steps)
    {
        if (params3.size() != 3) throw py::value_error("params3 = (lamb, zeta, phi0)");
        amb_handle* h = integrator.cast<Integrator&>().handle();
        double ms = 0;
        int rc;
        {
            py::gil_scoped_release release;
            rc = amb_surrogate_advance(s_, h, nsteps, params3.data(), euler_steps, &ms);
        }
        check(rc);
        return ms;
    }

    py::tuple audit(py::object main, py::object judge, const std::vector<double>& params3, int64_t euler_steps,
                    double tol)
    {
        if (params3.size() != 3) throw py::value_error("params3 = (lamb, zeta, phi0)");
        amb_handle* hm = main.cast<Integrator&>().handle();
        amb_handle* hj = judge.cast<Integrator&>().handle();
        double err = 0;
        int accepted = 0, rc;
        {
            py::gil_scoped_release release;
            rc = amb_surrogate_audit(s_, hm, hj, params3.data(), euler_steps, tol, &err, &accepted);
        }
        check(rc);
        return py::make_tuple(err, accepted != 0);
    }

    void audit_begin(py::object main, py::object judge, const std::vector<double>& params3, int64_t euler_steps)
    {
        if (params3.size() != 3) throw py::value_error("params3 = (lamb, zeta, phi0)");
        amb_handle* hm = main.cast<Integrator&>().handle();
        amb_handle* hj = judge.cast<Integrator&>().handle();
        int rc;
        {
            py::gil_scoped_release release;
            rc = amb_surrogate_audit_begin(s_, hm, hj, params3.data(), euler_steps);
        }
        check(rc);
        pending_ = py::make_tuple(main, judge);  // keep both integrators alive until audit_end
    }

    py::tuple audit_end(double tol)
    {
        double err = 0;
        int accepted = 0, rc;
        {
            py::gil_scoped_release release;
            rc = amb_surrogate_audit_end(s_, tol, &err, &accepted);
        }
        pending_ = py::none();
        check(rc);
        return py::make_tuple(err, accepted != 0);
    }

   private:
    py::object pending_ = py::none();
    amb_surrogate* s_ = nullptr;
    int64_t ny_, nx_, np_ = 0;
    int max_batch_;
};

// Batched ensemble over amb_batch_* (no reference counterpart).
class Ensemble {
   public:
    Ensemble(const std::vector<Field>& fields, const std::vector<Stencil>& stencils, const std::vector<Model>& models,
             const std::vector<uint64_t>& seeds, int precision)
    {
        const size_t n = fields.size();
        if (n == 0) throw py::value_error("need at least one realisation");
        if (stencils.size() != n || models.size() != n || seeds.size() != n)
            throw py::value_error("fields, stencils, models and seeds must have the same length");
        std::vector<amb_realisation> desc(n);
        std::vector<const double*> ptrs(n);
        for (size_t i = 0; i < n; ++i) {
            if (fields[i].ndim() != 2) throw py::value_error("every field must be a 2d array");
            desc[i].ny = fields[i].shape(0);
            desc[i].nx = fields[i].shape(1);
            desc[i].stencil = stencils[i].c_stencil();
            desc[i].model = models[i].c_model();
            desc[i].seed = seeds[i];
            ptrs[i] = fields[i].data();
            shapes_.emplace_back(desc[i].ny, desc[i].nx);
        }
        check(amb_batch_create(desc.data(), (int64_t)n, ptrs.data(), precision, &b_));
    }
    ~Ensemble() { amb_batch_destroy(b_); }
    Ensemble(const Ensemble&) = delete;
    Ensemble& operator=(const Ensemble&) = delete;

    size_t size() const { return shapes_.size(); }

    void run(int64_t nsteps)
    {
        int rc;
        {
            py::gil_scoped_release release;
            rc = amb_batch_run(b_, nsteps);
        }
        check(rc);
        timestep_ += nsteps;
    }

    double run_timed(int64_t nsteps)
    {
        int rc;
        double ms = 0;
        {
            py::gil_scoped_release release;
            rc = amb_batch_run_timed(b_, nsteps, &ms);
        }
        check(rc);
        timestep_ += nsteps;
        return ms;
    }

    py::array_t<double> field(int64_t i) const
    {
        if (i < 0 || i >= (int64_t)shapes_.size()) throw py::index_error("realisation index out of range");
        py::array_t<double> out({(py::ssize_t)shapes_[i].first, (py::ssize_t)shapes_[i].second});
        check(amb_batch_get_field(b_, i, out.mutable_data()));
        return out;
    }

    py::tuple mass(int64_t i) const
    {
        double s = 0;
        int64_t bad = 0;
        check(amb_batch_mass(b_, i, &s, &bad));
        return py::make_tuple(s, bad);
    }

    py::tuple last_mass() const
    {
        double s = 0;
        int64_t bad = 0;
        check(amb_batch_last_mass(b_, &s, &bad));
        return py::make_tuple(s, bad);
    }

    std::vector<int> tiles() const
    {
        std::vector<int> t(shapes_.size());
        check(amb_batch_get_tiles(b_, t.data()));
        return t;
    }

    void set_tiles(const std::vector<int>& t)
    {
        if (t.size() != shapes_.size()) throw py::value_error("one tile index per realisation");
        check(amb_batch_set_tiles(b_, t.data()));
    }

    void set_resident(bool on) { check(amb_batch_set_resident(b_, on ? 1 : 0)); }

    int64_t resident_count() const
    {
        int64_t n = 0;
        check(amb_batch_resident_count(b_, &n));
        return n;
    }

    std::vector<int> autotune(int reps, bool cached)
    {
        std::vector<int> t(shapes_.size());
        int rc;
        {
            py::gil_scoped_release release;
            rc = amb_batch_autotune(b_, cached ? -reps : reps, t.data());
        }
        check(rc);
        return t;
    }

    int64_t timestep() const { return timestep_; }

   private:
    amb_batch* b_ = nullptr;
    std::vector<std::pair<int64_t, int64_t>> shapes_;
    int64_t timestep_ = 0;
};

}  // namespace

PYBIND11_MODULE(integrator, m)
{
    m.doc() = "MI355X-native Active Model B+ integrator (HIP kernels behind a C ABI)";
    static py::exception<DeviceError> exc(m, "CudaException");
    m.attr("HipException") = exc;
    py::register_exception_translator([](std::exception_ptr p) {
        try {
            if (p) std::rethrow_exception(p);
        } catch (const DeviceError& e) {
            exc(e.what());
        }
    });

    py::class_<Model>(m, "Model")
        .def(py::init<>())
        .def(py::init([](double a, double b, double c, double kappa, double lamb, double zeta, double T) {
                 return Model{a, b, c, kappa, lamb, zeta, T};
             }),
             py::arg("a"), py::arg("b"), py::arg("c"), py::arg("kappa"), py::arg("lamb"), py::arg("zeta"),
             py::arg("T"))
        .def(py::init([](const Model& o) { return Model(o); }), py::arg("model"))
        .def_readwrite("a", &Model::a)
        .def_readwrite("b", &Model::b)
        .def_readwrite("c", &Model::c)
        .def_readwrite("kappa", &Model::kappa)
        .def_readwrite("lamb", &Model::lamb)
        .def_readwrite("zeta", &Model::zeta)
        .def_readwrite("T", &Model::T)
        .def("as_dict", &model_dict)
        .def("as_tuple", &Model::as_tuple)
        .def("__eq__", [](const Model& a, const Model& b) { return a == b; })
        .def("__str__", [](const Model& x) { return py::str(model_dict(x)); })
        .def("__repr__", [](const Model& x) { return py::str("<Model ") + py::str(model_dict(x)) + py::str(">"); });

    py::class_<Stencil>(m, "Stencil")
        .def(py::init<>())
        .def(py::init([](double dt, double dx, double dy) { return Stencil{dt, dx, dy}; }), py::arg("dt"),
             py::arg("dx"), py::arg("dy"))
        .def(py::init([](const Stencil& o) { return Stencil(o); }), py::arg("stencil"))
        .def_readwrite("dt", &Stencil::dt)
        .def_readwrite("dx", &Stencil::dx)
        .def_readwrite("dy", &Stencil::dy)
        .def("as_dict", &stencil_dict)
        .def("as_tuple", &Stencil::as_tuple)
        .def("__eq__", [](const Stencil& a, const Stencil& b) { return a == b; })
        .def("__str__", [](const Stencil& x) { return py::str(stencil_dict(x)); })
        .def("__repr__",
             [](const Stencil& x) { return py::str("<Stencil ") + py::str(stencil_dict(x)) + py::str(">"); });

    py::class_<Integrator>(m, "Integrator", py::dynamic_attr())
        .def(py::init<const Field&, const Stencil&, const Model&, py::object, int, int64_t, int>(), py::arg("field"),
             py::arg("stencil"), py::arg("model"), py::kw_only(), py::arg("seed") = py::none(),
             py::arg("precision") = 64, py::arg("timestep") = 0, py::arg("tile") = 0)
        .def_property_readonly("stencil", &Integrator::stencil)
        .def_property_readonly("model", &Integrator::model)
        .def_property_readonly("field", &Integrator::field)
        .def_property_readonly("chemical_potential", [](const Integrator& s) { return s.scalar(AMB_OBS_MU); })
        .def_property_readonly("passive_chemical_potential",
                               [](const Integrator& s) { return s.scalar(AMB_OBS_MU_PASSIVE); })
        .def_property_readonly("active_chemical_potential",
                               [](const Integrator& s) { return s.scalar(AMB_OBS_MU_ACTIVE); })
        .def_property_readonly("current", [](const Integrator& s) { return s.vector(AMB_OBS_J); })
        .def_property_readonly("deterministic_current",
                               [](const Integrator& s) { return s.vector(AMB_OBS_J_DETERMINISTIC); })
        .def_property_readonly("passive_current", [](const Integrator& s) { return s.vector(AMB_OBS_J_PASSIVE); })
        .def_property_readonly("active_current", [](const Integrator& s) { return s.vector(AMB_OBS_J_ACTIVE); })
        .def_property_readonly("lambda_current", [](const Integrator& s) { return s.vector(AMB_OBS_J_LAMBDA); })
        .def_property_readonly("circulating_current",
                               [](const Integrator& s) { return s.vector(AMB_OBS_J_CIRCULATING); })
        .def_property_readonly("random_current", [](const Integrator& s) { return s.vector(AMB_OBS_J_RANDOM); })
        .def_property_readonly("time", &Integrator::time)
        .def_property_readonly("timestep", &Integrator::timestep)
        .def_property_readonly("options", &Integrator::options)
        .def_property_readonly("seed", [](const Integrator& s) { return s.options()["seed"]; })
        .def_property_readonly("precision", [](const Integrator& s) { return s.options()["precision"]; })
        .def("mass", &Integrator::mass, "(FP64 sum of the field, number of non-finite sites)")
        .def("last_mass", &Integrator::last_mass,
             "(FP64 sum, number of non-finite sites) found by the last run, fused into its last step")
        .def("run", &Integrator::run, py::arg("nsteps"))
        .def("run_timed", &Integrator::run_timed, py::arg("nsteps"),
             "run(nsteps) returning the device milliseconds of the steps (HIP events on the integrator's stream)")
        .def("set_noise", &Integrator::set_noise, py::arg("noise"),
             "test only: explicit unscaled normals (steps, 2, ny, nx) for the next steps")
        .def("set_tile", &Integrator::set_tile, py::arg("index"))
        .def("autotune", &Integrator::autotune, py::arg("reps") = 8, py::arg("cached") = false,
             "time every compiled tile shape on a copy of the state and keep the fastest "
             "(cached: take the tuning table's entry when there is one)")
        .def("set_field", &Integrator::set_field, py::arg("field"), "replace the field (step counter unchanged)")
        .def("structure_factor_bins", &Integrator::structure_factor_bins, py::arg("edges"),
             "(sums, counts) of |phi_k|^2 per numpy.digitize(|k|, edges, right=True) index, on the device");

    py::class_<Surrogate>(m, "Surrogate")
        .def(py::init<int64_t, int64_t, int, uint64_t>(), py::arg("ny"), py::arg("nx"), py::arg("max_batch") = 16,
             py::arg("seed") = 0)
        .def_property_readonly("param_count", &Surrogate::param_count)
        .def("get_params", &Surrogate::get_params)
        .def("set_params", &Surrogate::set_params, py::arg("params"))
        .def("loss_grad", &Surrogate::loss_grad, py::arg("phi"), py::arg("params"), py::arg("target"))
        .def("set_dataset", &Surrogate::set_dataset, py::arg("phi_in"), py::arg("phi_out"), py::arg("params"))
        .def("train", &Surrogate::train, py::arg("iterations"), py::arg("batch"), py::arg("lr"), py::arg("clip") = 0.0)
        .def("set_train_noise", &Surrogate::set_train_noise, py::arg("rel"), py::arg("seed") = 0)
        .def("predict", &Surrogate::predict, py::arg("phi"), py::arg("params"))
        .def("advance", &Surrogate::advance, py::arg("integrator"), py::arg("nsteps"), py::arg("params3"),
             py::arg("euler_steps_per_step"))
        .def("audit", &Surrogate::audit, py::arg("main"), py::arg("judge"), py::arg("params3"),
             py::arg("euler_steps"), py::arg("tol"))
        .def("audit_begin", &Surrogate::audit_begin, py::arg("main"), py::arg("judge"), py::arg("params3"),
             py::arg("euler_steps"))
        .def("audit_end", &Surrogate::audit_end, py::arg("tol"));

    py::class_<Ensemble>(m, "Ensemble", py::dynamic_attr())
        .def(py::init<const std::vector<Field>&, const std::vector<Stencil>&, const std::vector<Model>&,
                      const std::vector<uint64_t>&, int>(),
             py::arg("fields"), py::arg("stencils"), py::arg("models"), py::arg("seeds"), py::arg("precision") = 32)
        .def("__len__", &Ensemble::size)
        .def("run", &Ensemble::run, py::arg("nsteps"))
        .def("run_timed", &Ensemble::run_timed, py::arg("nsteps"))
        .def("field", &Ensemble::field, py::arg("index"))
        .def("mass", &Ensemble::mass, py::arg("index"))
        .def("last_mass", &Ensemble::last_mass,
             "(FP64 sum over all realisations, number of non-finite sites) found by the last run")
        .def_property_readonly("tiles", &Ensemble::tiles)
        .def("set_tiles", &Ensemble::set_tiles, py::arg("tiles"))
        .def("set_resident", &Ensemble::set_resident, py::arg("on"),
             "keep members that fit in LDS resident across steps (default on)")
        .def_property_readonly("resident_count", &Ensemble::resident_count)
        .def("autotune", &Ensemble::autotune, py::arg("reps") = 4, py::arg("cached") = false,
             "time every tile shape per grid size and noise mode (cached: take the tuning table's entries first)")
        .def_property_readonly("timestep", &Ensemble::timestep);

    m.def("device_count", []() {
        int n = 0;
        amb_device_count(&n);
        return n;
    });
    m.def("tuning_set", [](int precision, int64_t ny, int64_t nx, bool noisy, int tile) {
        check(amb_tuning_set(precision, ny, nx, noisy ? 1 : 0, tile));
    }, py::arg("precision"), py::arg("ny"), py::arg("nx"), py::arg("noisy"), py::arg("tile"));
    m.def("tuning_get", [](int precision, int64_t ny, int64_t nx, bool noisy) {
        int t = -1;
        check(amb_tuning_get(precision, ny, nx, noisy ? 1 : 0, &t));
        return t;
    }, py::arg("precision"), py::arg("ny"), py::arg("nx"), py::arg("noisy"));
    m.def("tuning_entries", []() {
        int64_t n = 0;
        check(amb_tuning_count(&n));
        std::vector<std::tuple<int, int64_t, int64_t, bool, int>> out;
        for (int64_t i = 0; i < n; ++i) {
            int p = 0, noisy = 0, t = 0;
            int64_t ny = 0, nx = 0;
            check(amb_tuning_entry(i, &p, &ny, &nx, &noisy, &t));
            out.emplace_back(p, ny, nx, noisy != 0, t);
        }
        return out;
    }, "(precision, ny, nx, noisy, tile index) per entry of the tuning table");
    m.def("tuning_clear", []() { check(amb_tuning_clear()); });
    m.def("tile_shapes", [](int precision) {
        int n = 0;
        check(amb_tile_count(precision, &n));
        std::vector<std::tuple<int, int, int>> out;
        for (int i = 0; i < n; ++i) {
            int tw = 0, th = 0, nt = 0;
            check(amb_tile_shape(precision, i, &tw, &th, &nt));
            out.emplace_back(tw, th, nt);
        }
        return out;
    });
}
