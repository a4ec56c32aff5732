// d3q19_addon.cc -- thin Node-API binding of libd3q19 (include/d3q19.h).
//
// Exposes a `Lattice` class with the step / read-back idiom of the
// reference's own TypeScript solver class (packages/lbm-multiphase/web/src/solver.ts:45-302:
// seed, step substeps, read macroscopic fields):
//
//   const lat = new Lattice({nx, ny, nz, tau, storage, boundary, ...});
//   lat.setEquilibrium(rho | null, u);   // Float64Array / Float32Array
//   lat.step(n);
//   lat.readPopulations("f64" | "f32");  // Float64Array | Float32Array, (19,Nx,Ny,Nz)
//   lat.readMoments();                   // {rho: Float64Array, u: Float64Array}
//   lat.diagnostics();                   // {mass, momentum: [3], kineticEnergy}
//   lat.setMomentRanges(lo, hi);         // "me16" storage: moment quantization ranges
//   lat.momentRange(stored?);            // {lo, hi} of the 19 moments of f^n (stored: of the stored state)
//   lat.health();                        // {checkHealth, nonfiniteCells, rhoMin, rhoMax, uMaxLat}
//   lat.close();
//
// All arrays are host memory; the library stages them to the GPU.
#include <node_api.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/d3q19.h"

namespace {

#define NAPI_CALL(env, call)                                                  \
    do {                                                                      \
        if ((call) != napi_ok) {                                              \
            napi_throw_error((env), nullptr, "Node-API call failed: " #call); \
            return nullptr;                                                   \
        }                                                                     \
    } while (0)

napi_value throw_status(napi_env env, const char* fn, int st) {
    std::string msg = std::string(fn) + " failed (status " + std::to_string(st) + "): " + d3q19_last_error();
    napi_throw_error(env, nullptr, msg.c_str());
    return nullptr;
}

struct Wrapped {
    d3q19_lattice* lat = nullptr;
    int32_t nx = 0, ny = 0, nz = 0;
};

bool get_prop_double(napi_env env, napi_value obj, const char* key, double* out) {
    bool has = false;
    if (napi_has_named_property(env, obj, key, &has) != napi_ok || !has) return false;
    napi_value v;
    napi_get_named_property(env, obj, key, &v);
    napi_valuetype t;
    napi_typeof(env, v, &t);
    if (t != napi_number) return false;
    return napi_get_value_double(env, v, out) == napi_ok;
}

bool get_prop_string(napi_env env, napi_value obj, const char* key, std::string* out) {
    bool has = false;
    if (napi_has_named_property(env, obj, key, &has) != napi_ok || !has) return false;
    napi_value v;
    napi_get_named_property(env, obj, key, &v);
    char buf[64];
    size_t len = 0;
    if (napi_get_value_string_utf8(env, v, buf, sizeof(buf), &len) != napi_ok) return false;
    *out = std::string(buf, len);
    return true;
}

bool get_prop_vec3(napi_env env, napi_value obj, const char* key, double out[3]) {
    bool has = false;
    if (napi_has_named_property(env, obj, key, &has) != napi_ok || !has) return false;
    napi_value arr;
    napi_get_named_property(env, obj, key, &arr);
    for (uint32_t i = 0; i < 3; ++i) {
        napi_value e;
        if (napi_get_element(env, arr, i, &e) != napi_ok) return false;
        if (napi_get_value_double(env, e, &out[i]) != napi_ok) return false;
    }
    return true;
}

// Typed array -> (pointer, dtype code, element count); -1 dtype if unsupported.
bool typed_array(napi_env env, napi_value v, void** data, int* dtype, size_t* n) {
    bool is = false;
    napi_is_typedarray(env, v, &is);
    if (!is) return false;
    napi_typedarray_type t;
    napi_value ab;
    size_t off;
    napi_get_typedarray_info(env, v, &t, n, data, &ab, &off);
    if (t == napi_float64_array)
        *dtype = D3Q19_F64;
    else if (t == napi_float32_array)
        *dtype = D3Q19_F32;
    else
        return false;
    return true;
}

napi_value make_typed(napi_env env, int dtype, size_t n, void** data) {
    napi_value ab, ta;
    const size_t es = dtype == D3Q19_F64 ? 8 : 4;
    NAPI_CALL(env, napi_create_arraybuffer(env, n * es, data, &ab));
    NAPI_CALL(env, napi_create_typedarray(env, dtype == D3Q19_F64 ? napi_float64_array : napi_float32_array, n,
                                          ab, 0, &ta));
    return ta;
}

Wrapped* unwrap(napi_env env, napi_callback_info info, size_t* argc, napi_value* argv) {
    napi_value self;
    napi_get_cb_info(env, info, argc, argv, &self, nullptr);
    Wrapped* w = nullptr;
    napi_unwrap(env, self, (void**)&w);
    if (!w || !w->lat) {
        napi_throw_error(env, nullptr, "Lattice is closed");
        return nullptr;
    }
    return w;
}

void finalize(napi_env, void* data, void*) {
    Wrapped* w = (Wrapped*)data;
    if (w->lat) d3q19_destroy(w->lat);
    delete w;
}

napi_value Construct(napi_env env, napi_callback_info info) {
    size_t argc = 1;
    napi_value argv[1], self;
    NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, &self, nullptr));
    if (argc < 1) {
        napi_throw_type_error(env, nullptr, "new Lattice(config) needs a config object");
        return nullptr;
    }
    d3q19_config cfg;
    d3q19_config_default(&cfg);
    double v;
    if (get_prop_double(env, argv[0], "nx", &v)) cfg.nx = (int32_t)v;
    if (get_prop_double(env, argv[0], "ny", &v)) cfg.ny = (int32_t)v;
    if (get_prop_double(env, argv[0], "nz", &v)) cfg.nz = (int32_t)v;
    if (get_prop_double(env, argv[0], "tau", &v)) cfg.tau = v;
    if (get_prop_double(env, argv[0], "device", &v)) cfg.device = (int32_t)v;
    std::string s;
    if (get_prop_string(env, argv[0], "storage", &s)) {
        if (s == "f64") cfg.storage = D3Q19_F64;
        else if (s == "f32") cfg.storage = D3Q19_F32;
        else if (s == "f16") cfg.storage = D3Q19_F16;
        else if (s == "bf16") cfg.storage = D3Q19_BF16;
        else if (s == "bf16sr") cfg.storage = D3Q19_BF16SR;
        else if (s == "me16") cfg.storage = D3Q19_ME16;
        else if (s == "me16ref") cfg.storage = D3Q19_ME16_REF;
        else { napi_throw_range_error(env, nullptr, "storage must be f64|f32|f16|bf16|bf16sr|me16|me16ref"); return nullptr; }
    }
    if (get_prop_string(env, argv[0], "boundary", &s)) {
        if (s == "periodic") cfg.boundary = D3Q19_BC_PERIODIC;
        else if (s == "ywalls") cfg.boundary = D3Q19_BC_YWALLS;
        else if (s == "halfway") cfg.boundary = D3Q19_BC_HALFWAY;
        else { napi_throw_range_error(env, nullptr, "boundary must be periodic|ywalls|halfway"); return nullptr; }
    }
    get_prop_vec3(env, argv[0], "wallVelocityTop", cfg.wall_velocity_top);
    get_prop_vec3(env, argv[0], "wallVelocityBottom", cfg.wall_velocity_bottom);
    if (get_prop_vec3(env, argv[0], "force", cfg.force)) cfg.force_mode = D3Q19_FORCE_UNIFORM;
    if (get_prop_double(env, argv[0], "wallFaces", &v)) cfg.wall_faces = (uint32_t)v;
    static const char* faces[6] = {"-x", "+x", "-y", "+y", "-z", "+z"};
    bool has = false;
    if (napi_has_named_property(env, argv[0], "faceVelocity", &has) == napi_ok && has) {
        napi_value fv;
        napi_get_named_property(env, argv[0], "faceVelocity", &fv);
        for (int k = 0; k < 6; ++k) get_prop_vec3(env, fv, faces[k], cfg.face_velocity[k]);
    }
    d3q19_lattice* lat = nullptr;
    int st = d3q19_create(&cfg, nullptr, &lat);
    if (st) return throw_status(env, "d3q19_create", st);
    Wrapped* w = new Wrapped();
    w->lat = lat;
    w->nx = cfg.nx;
    w->ny = cfg.ny;
    w->nz = cfg.nz;
    NAPI_CALL(env, napi_wrap(env, self, w, finalize, nullptr, nullptr));
    return self;
}

size_t ncell(const Wrapped* w) { return (size_t)w->nx * w->ny * w->nz; }

napi_value SetEquilibrium(napi_env env, napi_callback_info info) {
    size_t argc = 2;
    napi_value argv[2];
    Wrapped* w = unwrap(env, info, &argc, argv);
    if (!w) return nullptr;
    void *pu = nullptr, *pr = nullptr;
    int du = -1, dr = -1;
    size_t nu = 0, nr = 0;
    if (argc < 2 || !typed_array(env, argv[1], &pu, &du, &nu) || nu != 3 * ncell(w)) {
        napi_throw_type_error(env, nullptr, "setEquilibrium(rho|null, u): u must be a Float64Array/Float32Array of 3*N");
        return nullptr;
    }
    napi_valuetype t;
    napi_typeof(env, argv[0], &t);
    if (t != napi_null && t != napi_undefined) {
        if (!typed_array(env, argv[0], &pr, &dr, &nr) || nr != ncell(w) || dr != du) {
            napi_throw_type_error(env, nullptr, "rho must match u's type and hold N values");
            return nullptr;
        }
    }
    int st = d3q19_set_equilibrium(w->lat, pr, pu, du, 0);
    if (st) return throw_status(env, "d3q19_set_equilibrium", st);
    return nullptr;
}

napi_value SetPopulations(napi_env env, napi_callback_info info) {
    size_t argc = 1;
    napi_value argv[1];
    Wrapped* w = unwrap(env, info, &argc, argv);
    if (!w) return nullptr;
    void* p = nullptr;
    int dt = -1;
    size_t n = 0;
    if (argc < 1 || !typed_array(env, argv[0], &p, &dt, &n) || n != 19 * ncell(w)) {
        napi_throw_type_error(env, nullptr, "setPopulations(f): f must be a Float64Array/Float32Array of 19*N");
        return nullptr;
    }
    int st = d3q19_set_populations(w->lat, p, dt, 0);
    if (st) return throw_status(env, "d3q19_set_populations", st);
    return nullptr;
}

napi_value Step(napi_env env, napi_callback_info info) {
    size_t argc = 1;
    napi_value argv[1];
    Wrapped* w = unwrap(env, info, &argc, argv);
    if (!w) return nullptr;
    int32_t n = 1;
    napi_valuetype t = napi_undefined;
    if (argc >= 1) napi_typeof(env, argv[0], &t);
    if (t != napi_undefined && napi_get_value_int32(env, argv[0], &n) != napi_ok) {
        napi_throw_type_error(env, nullptr, "step(n): n must be a number");
        return nullptr;
    }
    int st = d3q19_step(w->lat, n);
    if (st) return throw_status(env, "d3q19_step", st);
    st = d3q19_synchronize(w->lat);
    if (st) return throw_status(env, "d3q19_synchronize", st);
    return nullptr;
}

napi_value ReadPopulations(napi_env env, napi_callback_info info) {
    size_t argc = 1;
    napi_value argv[1];
    Wrapped* w = unwrap(env, info, &argc, argv);
    if (!w) return nullptr;
    int dt = D3Q19_F64;
    std::string s;
    if (argc >= 1) {
        char buf[8];
        size_t len = 0;
        if (napi_get_value_string_utf8(env, argv[0], buf, sizeof(buf), &len) == napi_ok && std::string(buf, len) == "f32")
            dt = D3Q19_F32;
    }
    void* data = nullptr;
    napi_value out = make_typed(env, dt, 19 * ncell(w), &data);
    if (!out) return nullptr;
    int st = d3q19_read_populations(w->lat, data, dt, 0);
    if (st) return throw_status(env, "d3q19_read_populations", st);
    return out;
}

napi_value ReadMoments(napi_env env, napi_callback_info info) {
    size_t argc = 0;
    Wrapped* w = unwrap(env, info, &argc, nullptr);
    if (!w) return nullptr;
    void *pr = nullptr, *pu = nullptr;
    napi_value rho = make_typed(env, D3Q19_F64, ncell(w), &pr);
    napi_value u = make_typed(env, D3Q19_F64, 3 * ncell(w), &pu);
    if (!rho || !u) return nullptr;
    int st = d3q19_read_moments(w->lat, pr, pu, D3Q19_F64, 0);
    if (st) return throw_status(env, "d3q19_read_moments", st);
    napi_value obj;
    NAPI_CALL(env, napi_create_object(env, &obj));
    NAPI_CALL(env, napi_set_named_property(env, obj, "rho", rho));
    NAPI_CALL(env, napi_set_named_property(env, obj, "u", u));
    return obj;
}

napi_value Diagnostics(napi_env env, napi_callback_info info) {
    size_t argc = 0;
    Wrapped* w = unwrap(env, info, &argc, nullptr);
    if (!w) return nullptr;
    double d[5];
    int st = d3q19_diagnostics(w->lat, d);
    if (st) return throw_status(env, "d3q19_diagnostics", st);
    napi_value obj, mass, ke, mom;
    NAPI_CALL(env, napi_create_object(env, &obj));
    NAPI_CALL(env, napi_create_double(env, d[0], &mass));
    NAPI_CALL(env, napi_create_double(env, d[4], &ke));
    NAPI_CALL(env, napi_create_array_with_length(env, 3, &mom));
    for (uint32_t a = 0; a < 3; ++a) {
        napi_value e;
        NAPI_CALL(env, napi_create_double(env, d[1 + a], &e));
        NAPI_CALL(env, napi_set_element(env, mom, a, e));
    }
    NAPI_CALL(env, napi_set_named_property(env, obj, "mass", mass));
    NAPI_CALL(env, napi_set_named_property(env, obj, "momentum", mom));
    NAPI_CALL(env, napi_set_named_property(env, obj, "kineticEnergy", ke));
    return obj;
}

// setMomentRanges(lo, hi): Float64Array(19) each, "me16" lattices before a state is loaded.
napi_value SetMomentRanges(napi_env env, napi_callback_info info) {
    size_t argc = 2;
    napi_value argv[2];
    Wrapped* w = unwrap(env, info, &argc, argv);
    if (!w) return nullptr;
    void *lo = nullptr, *hi = nullptr;
    int dl = -1, dh = -1;
    size_t nl = 0, nh = 0;
    if (argc < 2 || !typed_array(env, argv[0], &lo, &dl, &nl) || !typed_array(env, argv[1], &hi, &dh, &nh) ||
        dl != D3Q19_F64 || dh != D3Q19_F64 || nl != 19 || nh != 19) {
        napi_throw_type_error(env, nullptr, "setMomentRanges(lo, hi): two Float64Array of 19 values");
        return nullptr;
    }
    int st = d3q19_set_moment_ranges(w->lat, (const double*)lo, (const double*)hi);
    if (st) return throw_status(env, "d3q19_set_moment_ranges", st);
    return nullptr;
}

// momentRange(stored = false): {lo, hi} of the 19 moments of the current
// state f^n, or (stored) of the state as the lattice stores it.
napi_value MomentRange(napi_env env, napi_callback_info info) {
    size_t argc = 1;
    napi_value argv[1];
    Wrapped* w = unwrap(env, info, &argc, argv);
    if (!w) return nullptr;
    bool stored = false;
    if (argc >= 1) {
        napi_valuetype t;
        NAPI_CALL(env, napi_typeof(env, argv[0], &t));
        if (t == napi_boolean) {
            NAPI_CALL(env, napi_get_value_bool(env, argv[0], &stored));
        } else if (t != napi_undefined) {
            napi_throw_type_error(env, nullptr, "momentRange(stored): stored must be a boolean");
            return nullptr;
        }
    }
    void *plo = nullptr, *phi = nullptr;
    napi_value lo = make_typed(env, D3Q19_F64, 19, &plo);
    napi_value hi = make_typed(env, D3Q19_F64, 19, &phi);
    if (!lo || !hi) return nullptr;
    int st = stored ? d3q19_moment_range_stored(w->lat, (double*)plo, (double*)phi)
                    : d3q19_moment_range(w->lat, (double*)plo, (double*)phi);
    if (st) return throw_status(env, stored ? "d3q19_moment_range_stored" : "d3q19_moment_range", st);
    napi_value obj;
    NAPI_CALL(env, napi_create_object(env, &obj));
    NAPI_CALL(env, napi_set_named_property(env, obj, "lo", lo));
    NAPI_CALL(env, napi_set_named_property(env, obj, "hi", hi));
    return obj;
}

// health(): {checkHealth, nonfiniteCells, rhoMin, rhoMax, uMaxLat} of the state.
napi_value Health(napi_env env, napi_callback_info info) {
    size_t argc = 0;
    Wrapped* w = unwrap(env, info, &argc, nullptr);
    if (!w) return nullptr;
    double h[4];
    int st = d3q19_health(w->lat, h);
    if (st) return throw_status(env, "d3q19_health", st);
    const char* names[5] = {"checkHealth", "nonfiniteCells", "rhoMin", "rhoMax", "uMaxLat"};
    const double vals[5] = {h[0] == 0.0 ? 0.0 : 1.0, h[0], h[1], h[2], h[3]};
    napi_value obj;
    NAPI_CALL(env, napi_create_object(env, &obj));
    for (int k = 0; k < 5; ++k) {
        napi_value v;
        NAPI_CALL(env, napi_create_double(env, vals[k], &v));
        NAPI_CALL(env, napi_set_named_property(env, obj, names[k], v));
    }
    return obj;
}

napi_value StepCount(napi_env env, napi_callback_info info) {
    size_t argc = 0;
    Wrapped* w = unwrap(env, info, &argc, nullptr);
    if (!w) return nullptr;
    napi_value v;
    NAPI_CALL(env, napi_create_int64(env, d3q19_step_count(w->lat), &v));
    return v;
}

napi_value Close(napi_env env, napi_callback_info info) {
    napi_value self;
    size_t argc = 0;
    NAPI_CALL(env, napi_get_cb_info(env, info, &argc, nullptr, &self, nullptr));
    Wrapped* w = nullptr;
    napi_unwrap(env, self, (void**)&w);
    if (w && w->lat) {
        d3q19_destroy(w->lat);
        w->lat = nullptr;
    }
    return nullptr;
}

napi_value Version(napi_env env, napi_callback_info) {
    napi_value v;
    NAPI_CALL(env, napi_create_string_utf8(env, d3q19_version(), NAPI_AUTO_LENGTH, &v));
    return v;
}

napi_value Init(napi_env env, napi_value exports) {
    napi_property_descriptor methods[] = {
        {"setEquilibrium", nullptr, SetEquilibrium, nullptr, nullptr, nullptr, napi_default, nullptr},
        {"setPopulations", nullptr, SetPopulations, nullptr, nullptr, nullptr, napi_default, nullptr},
        {"step", nullptr, Step, nullptr, nullptr, nullptr, napi_default, nullptr},
        {"readPopulations", nullptr, ReadPopulations, nullptr, nullptr, nullptr, napi_default, nullptr},
        {"readMoments", nullptr, ReadMoments, nullptr, nullptr, nullptr, napi_default, nullptr},
        {"diagnostics", nullptr, Diagnostics, nullptr, nullptr, nullptr, napi_default, nullptr},
        {"stepCount", nullptr, StepCount, nullptr, nullptr, nullptr, napi_default, nullptr},
        {"setMomentRanges", nullptr, SetMomentRanges, nullptr, nullptr, nullptr, napi_default, nullptr},
        {"momentRange", nullptr, MomentRange, nullptr, nullptr, nullptr, napi_default, nullptr},
        {"health", nullptr, Health, nullptr, nullptr, nullptr, napi_default, nullptr},
        {"close", nullptr, Close, nullptr, nullptr, nullptr, napi_default, nullptr},
    };
    napi_value cls;
    NAPI_CALL(env, napi_define_class(env, "Lattice", NAPI_AUTO_LENGTH, Construct, nullptr,
                                     sizeof(methods) / sizeof(methods[0]), methods, &cls));
    NAPI_CALL(env, napi_set_named_property(env, exports, "Lattice", cls));
    napi_value ver;
    NAPI_CALL(env, napi_create_function(env, "version", NAPI_AUTO_LENGTH, Version, nullptr, &ver));
    NAPI_CALL(env, napi_set_named_property(env, exports, "version", ver));
    return exports;
}

}  // namespace

NAPI_MODULE(NODE_GYP_MODULE_NAME, Init)
