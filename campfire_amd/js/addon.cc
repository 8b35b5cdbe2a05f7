// addon.cc — N-API binding of the campfire C ABI (include/campfire_hip.h) for
// Node.js.  Plain N-API (no node-addon-api, no node-gyp): builds with g++
// against the Node headers and links libcampfire_hip.so.  fluid_hip.mjs wraps
// these functions in the reference's `FluidSim` class (js/fluid.js:26-356).
//
// Every negative status becomes a thrown JS Error carrying cf_last_error();
// CF_ENOTREADY from takeReadback returns null so the shim can poll.
#include <node_api.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "campfire_hip.h"

#define NAPI_OK(call)                                                  \
  do {                                                                 \
    if ((call) != napi_ok) {                                           \
      napi_throw_error(env, nullptr, "N-API call failed: " #call);     \
      return nullptr;                                                  \
    }                                                                  \
  } while (0)

static napi_value throw_cf(napi_env env, int rc, const char* what) {
  char msg[640];
  snprintf(msg, sizeof msg, "%s: %s (code %d)", what, cf_last_error(), rc);
  char code[16];
  snprintf(code, sizeof code, "%d", rc);
  napi_throw_error(env, code, msg);
  return nullptr;
}

static bool get_args(napi_env env, napi_callback_info info, size_t want, napi_value* argv) {
  size_t argc = want;
  if (napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr) != napi_ok || argc < want) {
    napi_throw_type_error(env, nullptr, "wrong number of arguments");
    return false;
  }
  return true;
}

static cf_sim* sim_of(napi_env env, napi_value v) {
  void* p = nullptr;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p) {
    napi_throw_type_error(env, nullptr, "expected a sim handle");
    return nullptr;
  }
  return (cf_sim*)p;
}

static napi_value undefined(napi_env env) {
  napi_value u;
  napi_get_undefined(env, &u);
  return u;
}

static napi_value num(napi_env env, double d) {
  napi_value v;
  napi_create_double(env, d, &v);
  return v;
}

// create(nx, ny, dx, dt) -> handle (js/fluid.js:27-160)
static napi_value Create(napi_env env, napi_callback_info info) {
  napi_value a[4];
  if (!get_args(env, info, 4, a)) return nullptr;
  int32_t nx, ny;
  double dx, dt;
  NAPI_OK(napi_get_value_int32(env, a[0], &nx));
  NAPI_OK(napi_get_value_int32(env, a[1], &ny));
  NAPI_OK(napi_get_value_double(env, a[2], &dx));
  NAPI_OK(napi_get_value_double(env, a[3], &dt));
  cf_sim* s = nullptr;
  int rc = cf_sim_create(nx, ny, dx, dt, &s);
  if (rc) return throw_cf(env, rc, rc == CF_EBADSIZE ? "coarsest level too large" : "init");
  napi_value h;
  NAPI_OK(napi_create_external(env, s, nullptr, nullptr, &h));
  return h;
}

static napi_value Destroy(napi_env env, napi_callback_info info) {
  napi_value a[1];
  if (!get_args(env, info, 1, a)) return nullptr;
  cf_sim* s = sim_of(env, a[0]);
  if (!s) return nullptr;
  cf_sim_destroy(s);
  return undefined(env);
}

// simple (handle) -> status calls
#define SIMPLE(Name, fn, what)                                        \
  static napi_value Name(napi_env env, napi_callback_info info) {     \
    napi_value a[1];                                                  \
    if (!get_args(env, info, 1, a)) return nullptr;                   \
    cf_sim* s = sim_of(env, a[0]);                                    \
    if (!s) return nullptr;                                           \
    int rc = fn(s);                                                   \
    if (rc < 0) return throw_cf(env, rc, what);                       \
    return undefined(env);                                            \
  }
SIMPLE(Reset, cf_sim_reset, "reset")
SIMPLE(WriteParams, cf_sim_write_params, "writeParams")
SIMPLE(ClearGeometry, cf_sim_clear_geometry, "clearGeometry")
SIMPLE(Geometry, cf_sim_geometry, "encodeGeometry")
SIMPLE(Radiation, cf_sim_radiation, "encodeRadiation")
SIMPLE(Synchronize, cf_sim_synchronize, "synchronize")
SIMPLE(Status, cf_sim_status, "status")

// setParam(h, key, value)  (js/fluid.js:178)
static napi_value SetParam(napi_env env, napi_callback_info info) {
  napi_value a[3];
  if (!get_args(env, info, 3, a)) return nullptr;
  cf_sim* s = sim_of(env, a[0]);
  if (!s) return nullptr;
  char key[64];
  size_t n;
  NAPI_OK(napi_get_value_string_utf8(env, a[1], key, sizeof key, &n));
  double v;
  NAPI_OK(napi_get_value_double(env, a[2], &v));
  int rc = cf_sim_set_param(s, key, v);
  if (rc < 0) return throw_cf(env, rc, "setParam");
  return undefined(env);
}

static napi_value GetParam(napi_env env, napi_callback_info info) {
  napi_value a[2];
  if (!get_args(env, info, 2, a)) return nullptr;
  cf_sim* s = sim_of(env, a[0]);
  if (!s) return nullptr;
  char key[64];
  size_t n;
  NAPI_OK(napi_get_value_string_utf8(env, a[1], key, sizeof key, &n));
  double v = 0;
  int rc = cf_sim_get_param(s, key, &v);
  if (rc < 0) return throw_cf(env, rc, "getParam");
  return num(env, v);
}

// uploadPieces(h, Float32Array(64*20), n)  (js/fluid.js:201-204)
static napi_value UploadPieces(napi_env env, napi_callback_info info) {
  napi_value a[3];
  if (!get_args(env, info, 3, a)) return nullptr;
  cf_sim* s = sim_of(env, a[0]);
  if (!s) return nullptr;
  napi_typedarray_type t;
  size_t len;
  void* data;
  napi_value ab;
  size_t off;
  if (napi_get_typedarray_info(env, a[1], &t, &len, &data, &ab, &off) != napi_ok || t != napi_float32_array ||
      len != CF_MAX_PIECES * CF_PIECE_FLOATS) {
    napi_throw_type_error(env, nullptr, "pieces must be a Float32Array(64*20)");
    return nullptr;
  }
  int32_t n;
  NAPI_OK(napi_get_value_int32(env, a[2], &n));
  int rc = cf_sim_upload_pieces(s, (const float*)data, n);
  if (rc < 0) return throw_cf(env, rc, "uploadPieces");
  return undefined(env);
}

// step(h, n)  (n x encodeStep, js/fluid.js:225-253)
static napi_value Step(napi_env env, napi_callback_info info) {
  napi_value a[2];
  if (!get_args(env, info, 2, a)) return nullptr;
  cf_sim* s = sim_of(env, a[0]);
  if (!s) return nullptr;
  int32_t n;
  NAPI_OK(napi_get_value_int32(env, a[1], &n));
  int rc = cf_sim_step(s, n);
  if (rc < 0) return throw_cf(env, rc, "encodeStep");
  return undefined(env);
}

// encodeReadback(h) -> slot or -1 when the ring slot is busy (js/fluid.js:315-324)
static napi_value EncodeReadback(napi_env env, napi_callback_info info) {
  napi_value a[1];
  if (!get_args(env, info, 1, a)) return nullptr;
  cf_sim* s = sim_of(env, a[0]);
  if (!s) return nullptr;
  int rc = cf_sim_encode_readback(s);
  if (rc == CF_EAGAIN) return num(env, -1);
  if (rc < 0) return throw_cf(env, rc, "encodeReadback");
  return num(env, rc);
}

static napi_value typed(napi_env env, napi_typedarray_type t, size_t elems, const void* src) {
  void* dst = nullptr;
  napi_value ab, arr;
  if (napi_create_arraybuffer(env, elems * 4, &dst, &ab) != napi_ok) return nullptr;
  memcpy(dst, src, elems * 4);
  if (napi_create_typedarray(env, t, elems, ab, 0, &arr) != napi_ok) return nullptr;
  return arr;
}

// takeReadback(h, slot, wait) -> {acc: Int32Array, rad: Float32Array, faceW: Int32Array} | null
static napi_value TakeReadback(napi_env env, napi_callback_info info) {
  napi_value a[3];
  if (!get_args(env, info, 3, a)) return nullptr;
  cf_sim* s = sim_of(env, a[0]);
  if (!s) return nullptr;
  int32_t slot;
  bool wait;
  NAPI_OK(napi_get_value_int32(env, a[1], &slot));
  NAPI_OK(napi_get_value_bool(env, a[2], &wait));
  static int32_t acc[CF_ACC_LEN], faceW[CF_MAX_PIECES * 4];
  static float rad[CF_RAD_LEN * 4];
  int rc = cf_sim_take_readback(s, slot, wait ? 1 : 0, acc, rad, faceW);
  if (rc == CF_ENOTREADY) {
    napi_value nul;
    napi_get_null(env, &nul);
    return nul;
  }
  if (rc < 0) return throw_cf(env, rc, "readback");
  napi_value o;
  NAPI_OK(napi_create_object(env, &o));
  NAPI_OK(napi_set_named_property(env, o, "acc", typed(env, napi_int32_array, CF_ACC_LEN, acc)));
  NAPI_OK(napi_set_named_property(env, o, "rad", typed(env, napi_float32_array, CF_RAD_LEN * 4, rad)));
  NAPI_OK(napi_set_named_property(env, o, "faceW", typed(env, napi_int32_array, CF_MAX_PIECES * 4, faceW)));
  return o;
}

// readField(h, name) -> ArrayBuffer copy of the buffer (js/fluid.js:345-355)
static napi_value ReadField(napi_env env, napi_callback_info info) {
  napi_value a[2];
  if (!get_args(env, info, 2, a)) return nullptr;
  cf_sim* s = sim_of(env, a[0]);
  if (!s) return nullptr;
  char name[32];
  size_t n;
  NAPI_OK(napi_get_value_string_utf8(env, a[1], name, sizeof name, &n));
  size_t bytes = 0;
  int rc = cf_sim_field_bytes(s, name, &bytes);
  if (rc < 0) return throw_cf(env, rc, "readBuffer");
  void* dst = nullptr;
  napi_value ab;
  NAPI_OK(napi_create_arraybuffer(env, bytes, &dst, &ab));
  rc = cf_sim_read_field(s, name, dst, bytes);
  if (rc < 0) return throw_cf(env, rc, "readBuffer");
  return ab;
}

// writeField(h, name, typedArray)
static napi_value WriteField(napi_env env, napi_callback_info info) {
  napi_value a[3];
  if (!get_args(env, info, 3, a)) return nullptr;
  cf_sim* s = sim_of(env, a[0]);
  if (!s) return nullptr;
  char name[32];
  size_t n;
  NAPI_OK(napi_get_value_string_utf8(env, a[1], name, sizeof name, &n));
  napi_typedarray_type t;
  size_t len, off;
  void* data;
  napi_value ab;
  if (napi_get_typedarray_info(env, a[2], &t, &len, &data, &ab, &off) != napi_ok ||
      (t != napi_float32_array && t != napi_int32_array && t != napi_uint32_array)) {
    napi_throw_type_error(env, nullptr, "expected a 32-bit typed array");
    return nullptr;
  }
  int rc = cf_sim_write_field(s, name, data, len * 4);
  if (rc < 0) return throw_cf(env, rc, "writeField");
  return undefined(env);
}

// info(h) -> {nx, ny, dx, dt, height, time, frame, diffSub, levels: [[nx, ny], ...]}
static napi_value Info(napi_env env, napi_callback_info info) {
  napi_value a[1];
  if (!get_args(env, info, 1, a)) return nullptr;
  cf_sim* s = sim_of(env, a[0]);
  if (!s) return nullptr;
  cf_sim_info in;
  int rc = cf_sim_get_info(s, &in);
  if (rc < 0) return throw_cf(env, rc, "info");
  napi_value o, lv;
  NAPI_OK(napi_create_object(env, &o));
  napi_set_named_property(env, o, "nx", num(env, in.nx));
  napi_set_named_property(env, o, "ny", num(env, in.ny));
  napi_set_named_property(env, o, "dx", num(env, in.dx));
  napi_set_named_property(env, o, "dt", num(env, in.dt));
  napi_set_named_property(env, o, "height", num(env, in.height));
  napi_set_named_property(env, o, "time", num(env, in.time));
  napi_set_named_property(env, o, "frame", num(env, in.frame));
  napi_set_named_property(env, o, "diffSub", num(env, in.diff_sub));
  NAPI_OK(napi_create_array_with_length(env, in.nlevels, &lv));
  for (int l = 0; l < in.nlevels; l++) {
    napi_value pair;
    napi_create_array_with_length(env, 2, &pair);
    napi_set_element(env, pair, 0, num(env, in.level_nx[l]));
    napi_set_element(env, pair, 1, num(env, in.level_ny[l]));
    napi_set_element(env, lv, l, pair);
  }
  napi_set_named_property(env, o, "levels", lv);
  return o;
}

// profileSteps(h, n) -> Float32Array(7) of summed ms per stage (js/fluid.js:256-292)
static napi_value ProfileSteps(napi_env env, napi_callback_info info) {
  napi_value a[2];
  if (!get_args(env, info, 2, a)) return nullptr;
  cf_sim* s = sim_of(env, a[0]);
  if (!s) return nullptr;
  int32_t n;
  NAPI_OK(napi_get_value_int32(env, a[1], &n));
  float ms[7];
  int rc = cf_sim_profile_steps(s, n, ms);
  if (rc < 0) return throw_cf(env, rc, "encodeStepProfiled");
  return typed(env, napi_float32_array, 7, ms);
}

static napi_value Version(napi_env env, napi_callback_info info) {
  napi_value v;
  napi_create_string_utf8(env, cf_version(), NAPI_AUTO_LENGTH, &v);
  return v;
}

static napi_value Init(napi_env env, napi_value exports) {
  napi_property_descriptor d[] = {
      {"create", 0, Create, 0, 0, 0, napi_default, 0},
      {"destroy", 0, Destroy, 0, 0, 0, napi_default, 0},
      {"reset", 0, Reset, 0, 0, 0, napi_default, 0},
      {"setParam", 0, SetParam, 0, 0, 0, napi_default, 0},
      {"getParam", 0, GetParam, 0, 0, 0, napi_default, 0},
      {"writeParams", 0, WriteParams, 0, 0, 0, napi_default, 0},
      {"uploadPieces", 0, UploadPieces, 0, 0, 0, napi_default, 0},
      {"clearGeometry", 0, ClearGeometry, 0, 0, 0, napi_default, 0},
      {"geometry", 0, Geometry, 0, 0, 0, napi_default, 0},
      {"step", 0, Step, 0, 0, 0, napi_default, 0},
      {"radiation", 0, Radiation, 0, 0, 0, napi_default, 0},
      {"encodeReadback", 0, EncodeReadback, 0, 0, 0, napi_default, 0},
      {"takeReadback", 0, TakeReadback, 0, 0, 0, napi_default, 0},
      {"readField", 0, ReadField, 0, 0, 0, napi_default, 0},
      {"writeField", 0, WriteField, 0, 0, 0, napi_default, 0},
      {"info", 0, Info, 0, 0, 0, napi_default, 0},
      {"profileSteps", 0, ProfileSteps, 0, 0, 0, napi_default, 0},
      {"synchronize", 0, Synchronize, 0, 0, 0, napi_default, 0},
      {"status", 0, Status, 0, 0, 0, napi_default, 0},
      {"version", 0, Version, 0, 0, 0, napi_default, 0},
  };
  napi_define_properties(env, exports, sizeof d / sizeof d[0], d);
  return exports;
}

NAPI_MODULE(NODE_GYP_MODULE_NAME, Init)
