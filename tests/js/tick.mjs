// Drives the GPU core through d3-force-wasm_amd/js/core.js exactly as the
// reference's facade does (src/simulation.js tick loop, src/manyBody.js,
// src/link.js and src/center.js force calls), for the end-to-end N-API test.
// With "radius" in the input, src/collide.js's force runs after the link
// force. Input: a JSON file {x, y, source, target, ticks[, radius]}; output:
// x, y, vx, vy of every node after the ticks, as 16-hex-digit float64 bit
// patterns.
import {readFileSync} from "fs";
import {context, Slot, MANY_BODY, LINK, CENTER, COLLIDE, THETA2, DISTANCE_MIN2, DISTANCE_MAX2, ITERATIONS, CX, CY,
        STRENGTH, F0, F1, F2, SOURCE, TARGET} from "../../d3-force-wasm_amd/js/core.js";

const inp = JSON.parse(readFileSync(process.argv[2], "utf8"));
const n = inp.x.length, m = inp.source.length;
const nodes = inp.x.map((x, i) => ({index: i, x, y: inp.y[i], vx: 0, vy: 0}));
let state = 1;
const random = () => (state = (1664525 * state + 1013904223) % 4294967296) / 4294967296;

const charge = new Slot(MANY_BODY);
charge.bind(nodes), charge.count(n), charge.f64(F0, n).fill(-30);
const count = new Array(n).fill(0);
for (let e = 0; e < m; ++e) count[inp.source[e]] += 1, count[inp.target[e]] += 1;
const link = new Slot(LINK);
link.bind(nodes), link.count(m);
const src = link.u32(SOURCE, m), tgt = link.u32(TARGET, m);
for (let e = 0; e < m; ++e) src[e] = inp.source[e], tgt[e] = inp.target[e];
const bias = link.f64(F2, m), strength = link.f64(F0, m), distance = link.f64(F1, m);
for (let e = 0; e < m; ++e) {
  const a = count[inp.source[e]], b = count[inp.target[e]];
  bias[e] = a / (a + b), strength[e] = 1 / Math.min(a, b), distance[e] = 30;
}
const center = new Slot(CENTER);
center.bind(nodes);
let collide = null;
if (inp.radius !== undefined) {
  collide = new Slot(COLLIDE);
  collide.bind(nodes), collide.count(n), collide.f64(F0, n).fill(inp.radius);
}

let alpha = 1;
const alphaDecay = 1 - Math.pow(0.001, 1 / 300);
const ctx = context(nodes);
for (let t = 0; t < inp.ticks; ++t) {
  ctx.copyIn(nodes);
  ctx.live = true;
  alpha += (0 - alpha) * alphaDecay;
  charge.scalar(THETA2, 0.81), charge.scalar(DISTANCE_MIN2, 1), charge.scalar(DISTANCE_MAX2, Infinity);
  charge.run(nodes, alpha, random);
  link.scalar(ITERATIONS, 1);
  link.run(nodes, alpha, random);
  if (collide) collide.scalar(STRENGTH, 1), collide.scalar(ITERATIONS, 1), collide.run(nodes, 0, random);
  center.scalar(CX, 0), center.scalar(CY, 0), center.scalar(STRENGTH, 1);
  center.run(nodes, 0);
  ctx.integrate(0.6);
  ctx.live = false;
  ctx.copyOut(nodes);
}
const f64 = new Float64Array(1), u64 = new BigUint64Array(f64.buffer);
const bits = (v) => (f64[0] = v, u64[0].toString(16).padStart(16, "0"));
process.stdout.write(JSON.stringify(["x", "y", "vx", "vy"].map((k) => nodes.map((d) => bits(d[k])))));
