// Node N-API addon over libd3fgpu.so: the binding the reference's JS facade
// uses in place of its wasm instance (src/core.js:24-30). Each export wraps
// one C entry point of include/d3fgpu.h one to one; arrays the wasm core
// exposed as offsets into its linear memory are handed to JS as external
// ArrayBuffers over the library's pinned host memory. Before a call that may
// free or move that memory (a resize, a count change, a destroy, a tree
// build) the addon detaches every buffer it handed out over it, so a view
// kept across such a call reads as empty instead of reaching freed memory
// (the wasm core's views are detached by memory growth in the same way).
// The random callback is the
// JS function given to set_random, called synchronously on the JS thread
// inside force_run, as the wasm import env.random was (core.js:11-14).
#include <node_api.h>

#include <cstdint>
#include <cstring>
#include <vector>

#include "d3fgpu.h"

namespace {

napi_env g_env = nullptr;
napi_ref g_random = nullptr;

#define CALL(expr)                                    \
  do {                                                \
    if ((expr) != napi_ok) {                          \
      napi_throw_error(env, nullptr, "napi: " #expr); \
      return nullptr;                                 \
    }                                                 \
  } while (0)

double random_trampoline(void*) {
  if (!g_env || !g_random) return 0.5;
  napi_value fn, global, result;
  double v = 0.5;
  if (napi_get_reference_value(g_env, g_random, &fn) != napi_ok) return v;
  if (napi_get_global(g_env, &global) != napi_ok) return v;
  if (napi_call_function(g_env, global, fn, 0, nullptr, &result) != napi_ok) return v;
  napi_get_value_double(g_env, result, &v);
  return v;
}

bool args(napi_env env, napi_callback_info info, size_t want, napi_value* argv) {
  size_t argc = want;
  if (napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr) != napi_ok) return false;
  if (argc < want) {
    napi_throw_type_error(env, nullptr, "too few arguments");
    return false;
  }
  return true;
}

uint32_t u32(napi_env env, napi_value v) {
  uint32_t r = 0;
  napi_get_value_uint32(env, v, &r);
  return r;
}

double f64(napi_env env, napi_value v) {
  double r = 0;
  napi_get_value_double(env, v, &r);
  return r;
}

napi_value num(napi_env env, double v) {
  napi_value r;
  napi_create_double(env, v, &r);
  return r;
}

// Buffers handed out, by owner: the simulation, and the force (0: the
// simulation's own arrays, TREE: its exported tree); weak references.
constexpr uint32_t TREE = 0xffffffffu;
struct Held {
  uint32_t sim, force;
  napi_ref ref;
};
std::vector<Held> g_held;

// Detaches the held buffers of simulation `sim` whose owner matches `force`
// (ALL: every owner of the simulation) and forgets them.
constexpr uint32_t ALL = 0xfffffffeu;
void detach(napi_env env, uint32_t sim, uint32_t force) {
  size_t keep = 0;
  for (size_t k = 0; k < g_held.size(); ++k) {
    Held& h = g_held[k];
    if (h.sim == sim && (force == ALL || h.force == force)) {
      napi_value ab;
      if (napi_get_reference_value(env, h.ref, &ab) == napi_ok && ab) napi_detach_arraybuffer(env, ab);
      napi_delete_reference(env, h.ref);
    } else {
      g_held[keep++] = h;
    }
  }
  g_held.resize(keep);
}

napi_value external(napi_env env, void* p, size_t bytes, uint32_t sim, uint32_t force) {
  napi_value r;
  if (!p || bytes == 0) {
    napi_get_null(env, &r);
    return r;
  }
  if (napi_create_external_arraybuffer(env, p, bytes, nullptr, nullptr, &r) != napi_ok) {
    napi_get_null(env, &r);
    return r;
  }
  napi_ref ref;
  if (napi_create_reference(env, r, 0, &ref) == napi_ok) g_held.push_back(Held{sim, force, ref});
  return r;
}

napi_value SetRandom(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!args(env, info, 1, argv)) return nullptr;
  if (g_random) napi_delete_reference(env, g_random);
  g_random = nullptr;
  napi_valuetype t;
  napi_typeof(env, argv[0], &t);
  if (t == napi_function) CALL(napi_create_reference(env, argv[0], 1, &g_random));
  g_env = env;
  core_set_random(&random_trampoline, nullptr);
  return nullptr;
}

napi_value SimCreate(napi_env env, napi_callback_info) { return num(env, sim_create()); }

#define WRAP1(Name, fn)                                      \
  napi_value Name(napi_env env, napi_callback_info info) {   \
    napi_value argv[1];                                      \
    if (!args(env, info, 1, argv)) return nullptr;           \
    return num(env, fn(u32(env, argv[0])));                  \
  }
#define WRAP2(Name, fn)                                                \
  napi_value Name(napi_env env, napi_callback_info info) {             \
    napi_value argv[2];                                                \
    if (!args(env, info, 2, argv)) return nullptr;                     \
    return num(env, fn(u32(env, argv[0]), u32(env, argv[1])));         \
  }

WRAP2(SimSync, sim_sync)
WRAP2(SimSnapshot, sim_snapshot)
WRAP2(ForceCreate, force_create)

napi_value SimDestroy(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!args(env, info, 1, argv)) return nullptr;
  const uint32_t h = u32(env, argv[0]);
  detach(env, h, ALL);
  return num(env, sim_destroy(h));
}

// A new node count may grow the simulation's arrays and its forces'.
napi_value SimSetNodeCount(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!args(env, info, 2, argv)) return nullptr;
  const uint32_t h = u32(env, argv[0]);
  detach(env, h, ALL);
  return num(env, sim_set_node_count(h, u32(env, argv[1])));
}

napi_value TreeBuild(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!args(env, info, 2, argv)) return nullptr;
  const uint32_t h = u32(env, argv[0]);
  detach(env, h, TREE);
  return num(env, tree_build(h, u32(env, argv[1])));
}

napi_value ForceDestroy(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!args(env, info, 2, argv)) return nullptr;
  const uint32_t h = u32(env, argv[0]), fh = u32(env, argv[1]);
  detach(env, h, fh);
  return num(env, force_destroy(h, fh));
}

napi_value TreeInfo(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!args(env, info, 2, argv)) return nullptr;
  return num(env, (double)tree_info(u32(env, argv[0]), u32(env, argv[1])));
}

napi_value SimIntegrate(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!args(env, info, 2, argv)) return nullptr;
  return num(env, sim_integrate(u32(env, argv[0]), f64(env, argv[1])));
}

// sim_buffer(h, which, bytes): the array sim_ptr(h, which) points at.
napi_value SimBuffer(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!args(env, info, 3, argv)) return nullptr;
  const uint32_t h = u32(env, argv[0]);
  return external(env, sim_ptr(h, u32(env, argv[1])), u32(env, argv[2]), h, 0);
}

// tree_buffer(h, which): 2 cells (len x 88 bytes), 3 extent (32), 4 perm (m x 4).
napi_value TreeBuffer(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!args(env, info, 2, argv)) return nullptr;
  const uint32_t h = u32(env, argv[0]), which = u32(env, argv[1]);
  size_t bytes = 0;
  if (which == 2) bytes = (size_t)tree_info(h, 0) * (size_t)tree_info(h, 5);
  else if (which == 3) bytes = 32;
  else if (which == 4) bytes = (size_t)tree_info(h, 6) * 4;
  return external(env, (void*)(uintptr_t)tree_info(h, which), bytes, h, TREE);
}

// tree_order(h, after) -> Uint32Array of the cell count.
napi_value TreeOrder(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!args(env, info, 2, argv)) return nullptr;
  const uint32_t h = u32(env, argv[0]);
  const size_t len = (size_t)tree_info(h, 0);
  void* data = nullptr;
  napi_value ab, arr;
  CALL(napi_create_arraybuffer(env, (len > 0 ? len : 1) * 4, &data, &ab));
  const int32_t st = tree_order(h, u32(env, argv[1]), (uint32_t*)data);
  if (st < 0) return num(env, st);
  CALL(napi_create_typedarray(env, napi_uint32_array, len, ab, 0, &arr));
  return arr;
}

napi_value ForceSetCount(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!args(env, info, 3, argv)) return nullptr;
  const uint32_t h = u32(env, argv[0]), fh = u32(env, argv[1]);
  detach(env, h, fh);
  return num(env, force_set_count(h, fh, u32(env, argv[2])));
}

// force_buffer(h, fh, which, bytes): the array force_ptr(h, fh, which) points at.
napi_value ForceBuffer(napi_env env, napi_callback_info info) {
  napi_value argv[4];
  if (!args(env, info, 4, argv)) return nullptr;
  const uint32_t h = u32(env, argv[0]), fh = u32(env, argv[1]);
  return external(env, force_ptr(h, fh, u32(env, argv[2])), u32(env, argv[3]), h, fh);
}

napi_value ForceSetScalar(napi_env env, napi_callback_info info) {
  napi_value argv[4];
  if (!args(env, info, 4, argv)) return nullptr;
  return num(env, force_set_scalar(u32(env, argv[0]), u32(env, argv[1]), u32(env, argv[2]), f64(env, argv[3])));
}

napi_value ForceGetScalar(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!args(env, info, 3, argv)) return nullptr;
  return num(env, force_get_scalar(u32(env, argv[0]), u32(env, argv[1]), u32(env, argv[2])));
}

napi_value ForceRun(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!args(env, info, 3, argv)) return nullptr;
  g_env = env;
  return num(env, force_run(u32(env, argv[0]), u32(env, argv[1]), f64(env, argv[2])));
}

napi_value Init(napi_env env, napi_value exports) {
  struct { const char* name; napi_callback cb; } fns[] = {
      {"set_random", SetRandom},       {"sim_create", SimCreate},         {"sim_destroy", SimDestroy},
      {"sim_set_node_count", SimSetNodeCount}, {"sim_sync", SimSync},     {"sim_snapshot", SimSnapshot},
      {"sim_integrate", SimIntegrate}, {"sim_buffer", SimBuffer},         {"tree_build", TreeBuild},
      {"tree_info", TreeInfo},         {"tree_buffer", TreeBuffer},       {"tree_order", TreeOrder},
      {"force_create", ForceCreate},   {"force_destroy", ForceDestroy},   {"force_set_count", ForceSetCount},
      {"force_buffer", ForceBuffer},   {"force_set_scalar", ForceSetScalar},
      {"force_get_scalar", ForceGetScalar}, {"force_run", ForceRun},
  };
  for (auto& f : fns) {
    napi_value fn;
    if (napi_create_function(env, f.name, NAPI_AUTO_LENGTH, f.cb, nullptr, &fn) != napi_ok) return nullptr;
    if (napi_set_named_property(env, exports, f.name, fn) != napi_ok) return nullptr;
  }
  return exports;
}

}  // namespace

NAPI_MODULE(NODE_GYP_MODULE_NAME, Init)
