// Loader and sync layer of the d3-force facade over the MI355X tick engine.
//
// A drop-in for the reference's src/core.js: the same exports (force kinds,
// scalar and array slots, check, context, Slot, slot, configure, testing),
// so src/simulation.js and the force facades run unchanged. The wasm
// instance is replaced by the N-API addon d3fgpu.node over libd3fgpu.so:
//  - typed-array views cover the library's pinned host arrays and are made
//    again after GROWN or a resize (as core.js:106-119 does on memory growth);
//  - copyIn ends with sim_sync(H2D) and copyOut starts with sim_sync(D2H),
//    because writes to the views cannot be intercepted;
//  - the running force's random source is installed with set_random before
//    each run and restored after it, as core.js:223-229 swaps `random`.
// There is no thread pool: every simulation runs on the GPU core.

import {createRequire} from "module";

const require = createRequire(import.meta.url);
const addon = require("./d3fgpu.node");

export const CENTER = 0, COLLIDE = 1, LINK = 2, MANY_BODY = 3, RADIAL = 4, X = 5, Y = 6;
export const STRENGTH = 0, CX = 1, CY = 2, THETA2 = 3, DISTANCE_MIN2 = 4, DISTANCE_MAX2 = 5, ITERATIONS = 6,
    PRECISION = 7, ORDER = 8, OPENING = 9, SEED = 10;
export const F0 = 0, F1 = 1, F2 = 2, SOURCE = 3, TARGET = 4;
export const SCALARS = 11;
export const FLOATS = [0, 1, 3, 1, 2, 2, 2];
export const OUT_OF_MEMORY = -2, WORKER_FAILED = -4;

const SYNC_H2D = 0, SYNC_D2H = 1;
const statuses = {"-1": "BAD_HANDLE", "-2": "OUT_OF_MEMORY", "-3": "BAD_ARGUMENT", "-4": "WORKER_FAILED"};

export function check(status) {
  if (status < 0) throw new Error(statuses[status] || "core status " + status);
  return status;
}

// The GPU core, in the shape simulation.js expects of `ctx.core`.
export const gpu = {exports: addon, threaded: false};
export const inlined = gpu;

// The running force's random source, called by the core's jiggle.
let random = null;
addon.set_random(() => random());

const contexts = new WeakMap();
export const registry = typeof FinalizationRegistry === "function"
  ? new FinalizationRegistry((h) => addon.sim_destroy(h))
  : null;

export const urls = {core: null, worker: null};
export function configure(options) {
  if (options && options.coreURL != null) urls.core = options.coreURL;
  if (options && options.workerURL != null) urls.worker = options.workerURL;
}

class Context {
  constructor() {
    this.core = gpu;
    if (!(this.h = addon.sim_create())) throw new Error("OUT_OF_MEMORY");
    this.n = 0;
    this.live = false;
    this.stale = true; // views must be made again
    this.precision = null;
    this.slots = new Set();
    this.inlinedOnly = true;
    if (registry) registry.register(this, this.h, this);
  }

  resize(n) {
    if (n !== this.n) check(addon.sim_set_node_count(this.h, n)), this.n = n, this.stale = true;
  }

  views() {
    if (!this.stale) return;
    const h = this.h, n = this.n, f64 = (w) => view(Float64Array, addon.sim_buffer(h, w, 8 * n), n);
    this.x = f64(0), this.y = f64(1), this.vx = f64(2), this.vy = f64(3), this.fx = f64(4), this.fy = f64(5);
    this.fixed = view(Uint8Array, addon.sim_buffer(h, 6, n), n);
    this.stale = false;
  }

  copyIn(nodes) {
    const n = nodes.length;
    this.resize(n), this.views();
    const {x, y, vx, vy, fx, fy, fixed} = this;
    for (let i = 0, d; i < n; ++i) {
      d = nodes[i];
      x[i] = d.x, y[i] = d.y, vx[i] = d.vx, vy[i] = d.vy;
      fx[i] = d.fx, fy[i] = d.fy;
      fixed[i] = (d.fx != null) | (d.fy != null) << 1;
    }
    check(addon.sim_sync(this.h, SYNC_H2D));
  }

  copyOut(nodes) {
    check(addon.sim_sync(this.h, SYNC_D2H));
    this.views();
    const {x, y, vx, vy} = this;
    for (let i = 0, n = this.n, d; i < n; ++i) {
      d = nodes[i];
      d.x = x[i], d.y = y[i], d.vx = vx[i], d.vy = vy[i];
    }
  }

  engage() {
    return 0;
  }

  step(slot, arg) {
    check(slot ? addon.force_run(this.h, slot.fh, arg) : addon.sim_integrate(this.h, arg));
  }

  integrate(velocityDecay) {
    this.step(null, velocityDecay);
  }
}

function view(Type, buffer, n) {
  return buffer ? new Type(buffer, 0, n) : new Type(0);
}

export function context(nodes) {
  let c = contexts.get(nodes);
  if (!c) contexts.set(nodes, c = new Context());
  return c;
}

export class Slot {
  constructor(kind) {
    this.kind = kind;
    this.ctx = null;
    this.fh = 0;
    this.m = 0;
  }

  bind(nodes) {
    const ctx = context(nodes);
    if (ctx !== this.ctx) {
      if (this.ctx) addon.force_destroy(this.ctx.h, this.fh), this.ctx.slots.delete(this);
      if (!(this.fh = addon.force_create(ctx.h, this.kind))) throw new Error("OUT_OF_MEMORY");
      this.ctx = ctx, this.m = 0;
      ctx.slots.add(this);
    }
    ctx.resize(nodes.length);
  }

  count(m) {
    check(addon.force_set_count(this.ctx.h, this.fh, m));
    this.m = m;
  }

  // A view of one parameter array; valid until the next core call.
  f64(which, m) {
    return view(Float64Array, addon.force_buffer(this.ctx.h, this.fh, which, 8 * m), m);
  }

  u32(which, m) {
    return view(Uint32Array, addon.force_buffer(this.ctx.h, this.fh, which, 4 * m), m);
  }

  scalar(which, v) {
    addon.force_set_scalar(this.ctx.h, this.fh, which, v);
  }

  run(nodes, alpha, rand) {
    const ctx = this.ctx, fast = ctx.precision;
    if (!ctx.live) ctx.copyIn(nodes);
    this.scalar(PRECISION, fast ? 1 : 0);
    if (fast) {
      this.scalar(ORDER, fast.order), this.scalar(OPENING, fast.opening);
      if (this.kind === MANY_BODY || this.kind === LINK || this.kind === COLLIDE) this.scalar(SEED, rand());
    }
    const previous = random;
    random = rand;
    try {
      ctx.step(this, alpha);
    } finally {
      random = previous;
      if (!ctx.live) ctx.copyOut(nodes);
    }
  }
}

export const slot = Symbol("slot");

export const testing = {
  get pool() { return null; },
  threshold() {},
  tune() {},
  request() { return Promise.resolve(); }
};
