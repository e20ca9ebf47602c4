// pygw.cpp -- the pybind11 module `GreensteinWinslow`, a drop-in for the
// reference binding (src/pyGW.cpp:160-344) on top of the C-ABI of
// include/gw_capi.h. Same classes, fields, defaults and run() signature; the
// integrator underneath is the HIP library (libgwhip.so).
//
// Differences from the reference binding, all deliberate:
//   - run() takes an optional keyword `seed`; without it the per-CRU streams
//     are seeded from std::random_device, as the reference's engines are.
//   - "mt19937" / "mt19937_64" select the xoshiro256** streams (the reference
//     seeds those engines from std::random_device, so no caller can depend on
//     their bit pattern); unknown names raise ValueError as before.
//   - GWCRUState(n) holds the seeded rest state (seed 0) instead of
//     uninitialised memory (src/pyGW.cpp:86).
//   - The record buffer has ceil(num_steps / record_every) rows (the reference
//     writes past num_steps / record_every rows when they do not divide).
//   - A non-positive Nernst argument raises RuntimeError instead of exit().
//   - run_reduced() (additive): run() with the records in the reduced encoding
//     of sim_dict_reduced_memory (src/serialization.py:11-39), formed on the
//     device.
#include <pybind11/functional.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <cstdlib>
#include <cstring>
#include <mutex>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "gw_capi.h"

namespace py = pybind11;
using namespace pybind11::literals;

namespace {

template <class T>
using Arr = py::array_t<T, py::array::c_style | py::array::forcecast>;

template <class T>
Arr<T> zeros(std::vector<py::ssize_t> shape)
{
    Arr<T> a(shape);
    std::memset(a.mutable_data(), 0, static_cast<size_t>(a.size()) * sizeof(T));
    return a;
}

// Convert `value` to an array of `dtype` with the given shape (-1 = any);
// a 1-D length-n array is accepted where the reference's Eigen row vector is.
template <class T>
Arr<T> coerce(const py::handle &value, const std::vector<py::ssize_t> &shape, const char *name)
{
    Arr<T> a = Arr<T>::ensure(value);
    if (!a) throw py::type_error(std::string(name) + ": expected an array");
    if (static_cast<size_t>(a.ndim()) != shape.size()) {
        // allow (1, n) for a length-n vector and vice versa
        if (shape.size() == 1 && a.ndim() == 2 && a.shape(0) == 1) {
            a = a.reshape({a.shape(1)});
        } else {
            throw py::value_error(std::string(name) + ": expected " + std::to_string(shape.size()) +
                                  "-D array, got " + std::to_string(a.ndim()) + "-D");
        }
    }
    for (size_t d = 0; d < shape.size(); ++d)
        if (shape[d] >= 0 && a.shape(d) != shape[d])
            throw py::value_error(std::string(name) + ": wrong size in dimension " + std::to_string(d));
    return a;
}

// ---- PyGlobalState (src/pyGW.cpp:31-76) ----------------------------------
struct PyGlobalState {
    double V = -91.382, Nai = 10.0, Ki = 131.84, Cai = 1.45273e-4, CaNSR = 0.908882;
    double CaLTRPN = 8.9282e-3;  // the binding's default differs from GW_utils.hpp:237 (8.98282e-3)
    double CaHTRPN = 0.137617, m = 5.33837e-4, h = 0.996345, j = 0.997315, xKs = 2.04171e-4;
    Arr<double> XKr, XKv14, XKv43;
    PyGlobalState()
    {
        gw_globals_t g;
        gw_globals_default(&g);
        XKr = Arr<double>({5});
        XKv14 = Arr<double>({10});
        XKv43 = Arr<double>({10});
        std::memcpy(XKr.mutable_data(), g.Kr, 5 * sizeof(double));
        std::memcpy(XKv14.mutable_data(), g.Kv14, 10 * sizeof(double));
        std::memcpy(XKv43.mutable_data(), g.Kv43, 10 * sizeof(double));
    }
    gw_globals_t to_c() const  // globals_from_python (src/pyGW.cpp:89-112)
    {
        gw_globals_t g;
        g.V = V; g.Nai = Nai; g.Ki = Ki; g.Cai = Cai; g.CaNSR = CaNSR; g.CaLTRPN = CaLTRPN; g.CaHTRPN = CaHTRPN;
        g.m = m; g.h = h; g.j = j; g.xKs = xKs;
        std::memcpy(g.Kr, XKr.data(), 5 * sizeof(double));
        std::memcpy(g.Kv14, XKv14.data(), 10 * sizeof(double));
        std::memcpy(g.Kv43, XKv43.data(), 10 * sizeof(double));
        return g;
    }
};

// ---- PyCRUState (src/pyGW.cpp:78-87) --------------------------------------
struct PyCRUState {
    int n;
    Arr<double> CaSS, CaJSR;
    Arr<int32_t> LCC, LCC_inactivation, RyR, ClCh;
    explicit PyCRUState(int nCRU) : n(nCRU)
    {
        if (nCRU <= 0) throw py::value_error("nCRU must be > 0");
        CaSS = Arr<double>({nCRU, 4});
        CaJSR = Arr<double>({nCRU});
        LCC = Arr<int32_t>({nCRU, 4});
        LCC_inactivation = Arr<int32_t>({nCRU, 4});
        RyR = Arr<int32_t>({nCRU, 4, 6});
        ClCh = Arr<int32_t>({nCRU, 4});
        gw_cru_state_t s{CaSS.mutable_data(), CaJSR.mutable_data(), LCC.mutable_data(), LCC_inactivation.mutable_data(),
                         RyR.mutable_data(), ClCh.mutable_data(), nullptr};
        gw_rest_state(nCRU, 0, GW_PRNG_XOSHIRO256SS, &s);
    }
};

// ---- PyGWSimulation (src/pyGW.hpp:22-53) ----------------------------------
struct PyGWSimulation {
    int nCRU;
    double tspan;
    Arr<double> t, V, m, h, j, Nai, Ki, Cai, CaNSR, CaLTRPN, CaHTRPN, xKs, XKr, XKv14, XKv43, CaJSR, CaSS;
    Arr<int32_t> LCC, LCC_inactivation, RyR, ClCh;
    PyGWSimulation(int nCRU_, int num_step, double tspan_) : nCRU(nCRU_), tspan(tspan_)
    {
        if (nCRU_ < 0 || num_step < 0) throw py::value_error("nCRU and num_step must be >= 0");
        const py::ssize_t R = num_step, N = nCRU_;
        for (Arr<double> *a : {&t, &V, &m, &h, &j, &Nai, &Ki, &Cai, &CaNSR, &CaLTRPN, &CaHTRPN, &xKs}) *a = zeros<double>({R});
        XKr = zeros<double>({R, 5});
        XKv14 = zeros<double>({R, 10});
        XKv43 = zeros<double>({R, 10});
        CaJSR = zeros<double>({R, N});
        CaSS = zeros<double>({R, N, 4});
        LCC = zeros<int32_t>({R, N, 4});
        LCC_inactivation = zeros<int32_t>({R, N, 4});
        ClCh = zeros<int32_t>({R, N, 4});
        RyR = zeros<int32_t>({R, N, 4, 6});
    }
};

std::mutex g_mutex;
gw_ctx *g_ctx = nullptr;

gw_ctx *context()
{
    if (g_ctx) return g_ctx;
    int device = 0;
    if (const char *s = std::getenv("GW_DEVICE")) device = std::atoi(s);
    char err[512] = {0};
    const int rc = gw_ctx_create(&g_ctx, device, err, sizeof(err));
    if (rc != GW_OK) throw std::runtime_error(std::string("GreensteinWinslow: no usable MI355X device: ") + err);
    return g_ctx;
}

void raise_status(int rc, const char *err)
{
    const std::string msg(err);
    if (rc == GW_E_INVALID_ARG) throw py::value_error(msg);
    throw std::runtime_error(msg);
}

int prng_id_from_name(const std::string &rng)
{
    // the eight names of run_PRNG_arg (src/pyGW.cpp:139-156)
    if (rng == "mt19937" || rng == "mt19937_64") return GW_PRNG_XOSHIRO256SS;
    const int id = gw_prng_from_name(rng.c_str());
    if (id < 0) throw std::invalid_argument(rng);
    return id;
}

// The arguments of run() as the C-ABI takes them.
struct RunInput {
    int prng;
    uint64_t seed;
    std::vector<double> istim;
    int64_t records;
    gw_globals_t g0;
    Arr<double> CaSS, CaJSR;
    Arr<int32_t> LCC, LCCi, RyR, ClCh;
    gw_cru_state_t c0;
};

RunInput run_input(int nCRU, double step_size, int num_steps, const py::object &Istim, int record_every,
                   PyCRUState &init_crus, PyGlobalState &init_globals, const std::string &rng, py::object seed)
{
    RunInput in;
    in.prng = prng_id_from_name(rng);
    if (nCRU <= 0) throw py::value_error("nCRU must be > 0");
    if (record_every <= 0) throw py::value_error("record_every must be > 0");
    if (num_steps < 0) throw py::value_error("num_steps must be >= 0");
    if (init_crus.CaSS.shape(0) != nCRU || init_crus.CaJSR.shape(0) != nCRU || init_crus.LCC.shape(0) != nCRU ||
        init_crus.LCC_inactivation.shape(0) != nCRU || init_crus.RyR.shape(0) != nCRU || init_crus.ClCh.shape(0) != nCRU)
        throw py::value_error("init_crus must hold nCRU CRUs");
    if (seed.is_none()) {
        std::random_device rd;
        in.seed = (static_cast<uint64_t>(rd()) << 32) ^ rd();
    } else {
        in.seed = seed.cast<uint64_t>();
    }
    // Istim is evaluated at the accumulated time before every step (src/pyGW.hpp:110-116)
    in.istim.resize(static_cast<size_t>(num_steps));
    {
        double t = 0.0;
        for (int i = 0; i < num_steps; ++i) {
            in.istim[i] = Istim(t).cast<double>();
            t += step_size;
        }
    }
    in.records = num_steps == 0 ? 0 : (static_cast<int64_t>(num_steps) + record_every - 1) / record_every;
    in.g0 = init_globals.to_c();
    in.CaSS = init_crus.CaSS; in.CaJSR = init_crus.CaJSR; in.LCC = init_crus.LCC;
    in.LCCi = init_crus.LCC_inactivation; in.RyR = init_crus.RyR; in.ClCh = init_crus.ClCh;
    in.c0 = gw_cru_state_t{const_cast<double *>(in.CaSS.data()), const_cast<double *>(in.CaJSR.data()),
                           const_cast<int32_t *>(in.LCC.data()), const_cast<int32_t *>(in.LCCi.data()),
                           const_cast<int32_t *>(in.RyR.data()), const_cast<int32_t *>(in.ClCh.data()), nullptr};
    return in;
}

// The precision keyword of run / run_reduced (additive; the reference runs FP64 only).
int32_t precision_from_name(const std::string &name)
{
    if (name == "fp64") return GW_PREC_FP64;
    if (name == "fp32_local") return GW_PREC_FP32_LOCAL;
    throw py::value_error("precision must be \"fp64\" or \"fp32_local\" (got \"" + name + "\")");
}

// run_PRNG_arg + run<PRNG> (src/pyGW.cpp:134-157, src/pyGW.hpp:102-122)
PyGWSimulation run(const gw_params_t &params, int nCRU, double step_size, int num_steps, const py::object &Istim,
                   int record_every, PyCRUState &init_crus, PyGlobalState &init_globals, const std::string &rng,
                   py::object seed, const std::string &precision)
{
    const int32_t prec = precision_from_name(precision);
    RunInput in = run_input(nCRU, step_size, num_steps, Istim, record_every, init_crus, init_globals, rng, seed);
    const int R = static_cast<int>(in.records);
    PyGWSimulation out(nCRU, R, num_steps * step_size);
    gw_record_t rec{};
    rec.num_records = R;
    rec.t = out.t.mutable_data(); rec.V = out.V.mutable_data(); rec.m = out.m.mutable_data();
    rec.h = out.h.mutable_data(); rec.j = out.j.mutable_data(); rec.Nai = out.Nai.mutable_data();
    rec.Ki = out.Ki.mutable_data(); rec.Cai = out.Cai.mutable_data(); rec.CaNSR = out.CaNSR.mutable_data();
    rec.CaLTRPN = out.CaLTRPN.mutable_data(); rec.CaHTRPN = out.CaHTRPN.mutable_data(); rec.xKs = out.xKs.mutable_data();
    rec.XKr = out.XKr.mutable_data(); rec.XKv14 = out.XKv14.mutable_data(); rec.XKv43 = out.XKv43.mutable_data();
    rec.CaJSR = out.CaJSR.mutable_data(); rec.CaSS = out.CaSS.mutable_data(); rec.LCC = out.LCC.mutable_data();
    rec.LCC_inactivation = out.LCC_inactivation.mutable_data(); rec.ClCh = out.ClCh.mutable_data();
    rec.RyR = out.RyR.mutable_data();
    char err[1024] = {0};
    int rc;
    {
        py::gil_scoped_release release;
        std::lock_guard<std::mutex> lock(g_mutex);
        gw_ctx *ctx = context();
        rc = gw_run(ctx, &params, nCRU, step_size, num_steps, in.istim.data(), record_every, &in.g0, &in.c0, in.seed,
                    in.prng, prec, &rec, err, sizeof(err));
    }
    if (rc != GW_OK) raise_status(rc, err);
    return out;
}

// run() with its records in the reduced encoding of sim_dict_reduced_memory
// (src/serialization.py:11-39), formed on the device: a dict with the same
// keys, dtypes and shapes as sim_dict_reduced_memory(GWSolution(run(...))).
py::dict run_reduced(const gw_params_t &params, int nCRU, double step_size, int num_steps, const py::object &Istim,
                     int record_every, PyCRUState &init_crus, PyGlobalState &init_globals, const std::string &rng,
                     py::object seed, const std::string &precision)
{
    const int32_t prec = precision_from_name(precision);
    RunInput in = run_input(nCRU, step_size, num_steps, Istim, record_every, init_crus, init_globals, rng, seed);
    const py::ssize_t R = in.records, N = nCRU;
    py::dict out;
    gw_record_reduced_t rec{};
    rec.num_records = R;
    const char *scalars[] = {"t", "V", "Nai", "Ki", "Cai", "CaNSR", "CaLTRPN", "CaHTRPN", "m", "h", "j", "xKs"};
    float **fields[] = {&rec.t, &rec.V, &rec.Nai, &rec.Ki, &rec.Cai, &rec.CaNSR, &rec.CaLTRPN, &rec.CaHTRPN,
                        &rec.m, &rec.h, &rec.j, &rec.xKs};
    for (int k = 0; k < 12; ++k) {
        auto a = zeros<float>({R});
        *fields[k] = a.mutable_data();
        out[scalars[k]] = a;
    }
    auto XKr = zeros<float>({R, 4}), XKv14 = zeros<float>({R, 9}), XKv43 = zeros<float>({R, 9});
    auto CaSS = zeros<float>({R, N, 4});
    auto CaJSR = zeros<float>({R, N});
    auto LCC = zeros<int8_t>({R, N, 4});
    auto LCCi = zeros<bool>({R, N, 4}), ClCh = zeros<bool>({R, N, 4});
    auto RyR = zeros<int16_t>({R, N, 4});
    rec.XKr = XKr.mutable_data(); rec.XKv14 = XKv14.mutable_data(); rec.XKv43 = XKv43.mutable_data();
    rec.CaSS = CaSS.mutable_data(); rec.CaJSR = CaJSR.mutable_data(); rec.LCC = LCC.mutable_data();
    rec.LCC_inactivation = reinterpret_cast<uint8_t *>(LCCi.mutable_data());
    rec.ClCh = reinterpret_cast<uint8_t *>(ClCh.mutable_data());
    rec.RyR = RyR.mutable_data();
    out["XKr"] = XKr; out["XKv14"] = XKv14; out["XKv43"] = XKv43; out["CaSS"] = CaSS; out["CaJSR"] = CaJSR;
    out["LCC"] = LCC; out["LCC_inactivation"] = LCCi; out["RyR"] = RyR; out["ClCh"] = ClCh;
    char err[1024] = {0};
    int rc;
    {
        py::gil_scoped_release release;
        std::lock_guard<std::mutex> lock(g_mutex);
        gw_ctx *ctx = context();
        rc = gw_run_reduced(ctx, &params, nCRU, step_size, num_steps, in.istim.data(), record_every, &in.g0, &in.c0,
                            in.seed, in.prng, prec, &rec, err, sizeof(err));
    }
    if (rc != GW_OK) raise_status(rc, err);
    return out;
}

}  // namespace

#define GW_DEF_ARR(cls, name, T, ...)                                                                    \
    .def_property(#name, [](cls &s) { return s.name; },                                                 \
                  [](cls &s, py::handle v) { s.name = coerce<T>(v, {__VA_ARGS__}, #name); })

PYBIND11_MODULE(GreensteinWinslow, m)
{
    m.doc() = "Greenstein-Winslow coupled integrator on MI355X (drop-in for the reference binding).";

    auto params = py::class_<gw_params_t>(m, "GWParameters");
    params.def(py::init([]() {
              gw_params_t p;
              gw_params_default(&p);
              return p;
          }))
        .def("__repr__", [](const gw_params_t &) { return "Greenstein and Winslow Model Parameters"; });
#define GW_BIND_FIELD_(type, name, dflt) \
    if (std::strncmp(#name, "_pad", 4) != 0) params.def_readwrite(#name, &gw_params_t::name);
    GW_PARAMETER_FIELDS(GW_BIND_FIELD_)
#undef GW_BIND_FIELD_

    py::class_<PyGWSimulation>(m, "GWVariables")
        .def(py::init<int, int, double>())
        GW_DEF_ARR(PyGWSimulation, t, double, -1)
        GW_DEF_ARR(PyGWSimulation, V, double, -1)
        GW_DEF_ARR(PyGWSimulation, m, double, -1)
        GW_DEF_ARR(PyGWSimulation, h, double, -1)
        GW_DEF_ARR(PyGWSimulation, j, double, -1)
        GW_DEF_ARR(PyGWSimulation, Nai, double, -1)
        GW_DEF_ARR(PyGWSimulation, Ki, double, -1)
        GW_DEF_ARR(PyGWSimulation, Cai, double, -1)
        GW_DEF_ARR(PyGWSimulation, CaNSR, double, -1)
        GW_DEF_ARR(PyGWSimulation, CaLTRPN, double, -1)
        GW_DEF_ARR(PyGWSimulation, CaHTRPN, double, -1)
        GW_DEF_ARR(PyGWSimulation, xKs, double, -1)
        GW_DEF_ARR(PyGWSimulation, XKr, double, -1, -1)
        GW_DEF_ARR(PyGWSimulation, XKv14, double, -1, -1)
        GW_DEF_ARR(PyGWSimulation, XKv43, double, -1, -1)
        GW_DEF_ARR(PyGWSimulation, CaJSR, double, -1, -1)
        GW_DEF_ARR(PyGWSimulation, CaSS, double, -1, -1, -1)
        GW_DEF_ARR(PyGWSimulation, LCC, int32_t, -1, -1, -1)
        GW_DEF_ARR(PyGWSimulation, LCC_inactivation, int32_t, -1, -1, -1)
        GW_DEF_ARR(PyGWSimulation, RyR, int32_t, -1, -1, -1, -1)
        GW_DEF_ARR(PyGWSimulation, ClCh, int32_t, -1, -1, -1)
        .def("__repr__", [](const PyGWSimulation &x) {
            return "Greenstein and Winslow model solution over the interval [0, " + std::to_string(x.tspan) +
                   "] with " + std::to_string(x.nCRU) + " CRUs";
        });

    py::class_<PyGlobalState>(m, "GWGlobalState")
        .def(py::init<>())
        .def_readwrite("V", &PyGlobalState::V)
        .def_readwrite("Nai", &PyGlobalState::Nai)
        .def_readwrite("Ki", &PyGlobalState::Ki)
        .def_readwrite("Cai", &PyGlobalState::Cai)
        .def_readwrite("CaNSR", &PyGlobalState::CaNSR)
        .def_readwrite("CaLTRPN", &PyGlobalState::CaLTRPN)
        .def_readwrite("CaHTRPN", &PyGlobalState::CaHTRPN)
        .def_readwrite("m", &PyGlobalState::m)
        .def_readwrite("h", &PyGlobalState::h)
        .def_readwrite("j", &PyGlobalState::j)
        .def_readwrite("xKs", &PyGlobalState::xKs)
        GW_DEF_ARR(PyGlobalState, XKr, double, 5)
        GW_DEF_ARR(PyGlobalState, XKv14, double, 10)
        GW_DEF_ARR(PyGlobalState, XKv43, double, 10);

    py::class_<PyCRUState>(m, "GWCRUState")
        .def(py::init<int>())
        .def_property("CaSS", [](PyCRUState &s) { return s.CaSS; },
                      [](PyCRUState &s, py::handle v) { s.CaSS = coerce<double>(v, {-1, 4}, "CaSS"); })
        .def_property("CaJSR", [](PyCRUState &s) { return s.CaJSR; },
                      [](PyCRUState &s, py::handle v) { s.CaJSR = coerce<double>(v, {-1}, "CaJSR"); })
        .def_property("LCC", [](PyCRUState &s) { return s.LCC; },
                      [](PyCRUState &s, py::handle v) { s.LCC = coerce<int32_t>(v, {-1, 4}, "LCC"); })
        .def_property("LCC_inactivation", [](PyCRUState &s) { return s.LCC_inactivation; },
                      [](PyCRUState &s, py::handle v) { s.LCC_inactivation = coerce<int32_t>(v, {-1, 4}, "LCC_inactivation"); })
        .def_property("RyR", [](PyCRUState &s) { return s.RyR; },
                      [](PyCRUState &s, py::handle v) { s.RyR = coerce<int32_t>(v, {-1, 4, 6}, "RyR"); })
        .def_property("ClCh", [](PyCRUState &s) { return s.ClCh; },
                      [](PyCRUState &s, py::handle v) { s.ClCh = coerce<int32_t>(v, {-1, 4}, "ClCh"); });

    m.def("run", &run, "Simulate the model", "parameters"_a, "nCRU"_a, "step_size"_a, "num_steps"_a, "Istim"_a,
          "record_every"_a, "init_crus"_a, "init_globals"_a = PyGlobalState(), "PRNG"_a = "mt19937_64",
          "seed"_a = py::none(), "precision"_a = "fp64");

    m.def("run_reduced", &run_reduced,
          "run() with its records in the reduced encoding of sim_dict_reduced_memory, formed on the device",
          "parameters"_a, "nCRU"_a, "step_size"_a, "num_steps"_a, "Istim"_a, "record_every"_a, "init_crus"_a,
          "init_globals"_a = PyGlobalState(), "PRNG"_a = "mt19937_64", "seed"_a = py::none(), "precision"_a = "fp64");

    m.def("rest_state", [](int nCRU, uint64_t seed, const std::string &rng) {
        PyCRUState s(nCRU);
        gw_cru_state_t c{s.CaSS.mutable_data(), s.CaJSR.mutable_data(), s.LCC.mutable_data(),
                         s.LCC_inactivation.mutable_data(), s.RyR.mutable_data(), s.ClCh.mutable_data(), nullptr};
        gw_rest_state(nCRU, seed, prng_id_from_name(rng), &c);
        return s;
    }, "Seeded rest state of nCRU CRUs (GW_utils.tpp:5-64 with per-CRU streams)", "nCRU"_a, "seed"_a = 0,
       "PRNG"_a = "xoshiro256**");

    m.add_object("_cleanup", py::capsule([]() {
        if (g_ctx) {
            gw_ctx_destroy(g_ctx);
            g_ctx = nullptr;
        }
    }));
}
