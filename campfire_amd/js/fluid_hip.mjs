// fluid_hip.mjs — the reference's FluidSim interface (js/fluid.js) over the
// MI355X gas solver (campfire_addon.node -> libcampfire_hip.so).
//
// A drop-in for js/fluid.js at its call sites (js/main.js:174-182,190-208,
// 247-285; dev/gas_test.html:37-87): same exports, constructor, method names,
// argument meaning and error behaviour.  `device`, `encoder` and `pass`
// arguments are accepted and ignored: the work is enqueued on the sim's own
// HIP stream in call order, which is the ordering WebGPU gives a single
// compute pass.  Written for Node >= 12 (no ?. or ??).
import { createRequire } from 'module';

const require = createRequire(import.meta.url);
const addon = require('./campfire_addon.node');

export const MAX_PIECES = 64;
export const PIECE_FLOATS = 20;
export const ACC_GLOBAL = 512;
export const ACC_LEN = ACC_GLOBAL + 16;
export const E_SCALE = 1e6, M_SCALE = 1e12, G_SCALE = 1e3;
export const N_FLOOR = 64;
export const RAD_LEN = MAX_PIECES * 4 + N_FLOOR;
export const SCENE_WIDTH = 0.7;

const STAGES = ['advect', 'diffuse', 'sources', 'forces', 'coef/rhs', 'vcycle', 'project'];

export class FluidSim {
  // new FluidSim(device, mode)  (js/fluid.js:27-47); mode.dx is optional and
  // defaults to SCENE_WIDTH / nx as in the reference
  constructor(device, mode) {
    this.device = device;
    this.nx = mode.nx; this.ny = mode.ny;
    this.dx = mode.dx !== undefined ? mode.dx : SCENE_WIDTH / mode.nx;
    this.height = this.dx * mode.ny;
    this.dt = mode.dt;
    this.h = null;
    this.pendingStage = -1;
  }

  async init() {  // js/fluid.js:49-160; throws 'coarsest level too large'
    this.h = addon.create(this.nx, this.ny, this.dx, this.dt);
    const inf = addon.info(this.h);
    this.levels = inf.levels.map(([nx, ny], l) => ({ nx, ny, h: this.dx * Math.pow(2, l) }));
    const self = this;
    // buf: named handles of the device buffers, readable through readBuffer
    this.buf = new Proxy({}, { get: (t, name) => (typeof name === 'string' ? { name, sim: self } : undefined) });
  }

  destroy() { if (this.h) { addon.destroy(this.h); this.h = null; } }

  get time() { return addon.info(this.h).time; }
  set time(v) { addon.setParam(this.h, 'time', v); }
  get frame() { return addon.info(this.h).frame; }
  set frame(v) { addon.setParam(this.h, 'frame', v); }
  get diffSub() { return addon.info(this.h).diffSub; }
  get params() {
    const names = ['Tamb', 'rhoAmb', 'YOamb', 'grav', 'cp', 'Pr', 'viscMult', 'diffMult', 'reactB', 'reactTa',
      'dHc', 'sO2', 'sootFormA', 'sootFormTa', 'sootOxA', 'sootOxTa', 'kappaSootC', 'kappaGas', 'exchAlpha',
      'depth', 'charA', 'charTa', 'nPieces', 'heaterX0', 'heaterX1', 'heaterY1', 'heaterT', 'burnerX0',
      'burnerX1', 'burnerFlux', 'burnerT', 'sootDensity', 'sigma', 'charO2'];
    const o = {};
    for (const k of names) o[k] = addon.getParam(this.h, k);
    return o;
  }

  reset() { addon.reset(this.h); }                                   // js/fluid.js:162-176
  setParam(k, v) { addon.setParam(this.h, k, v); }                   // js/fluid.js:178
  writeParams() { addon.writeParams(this.h); }                       // js/fluid.js:180-199
  uploadPieces(data, n) { addon.uploadPieces(this.h, data, n); }     // js/fluid.js:201-204
  clearGeometry(encoder) { addon.clearGeometry(this.h); }            // js/fluid.js:213
  encodeGeometry(pass) { addon.geometry(this.h); }                   // js/fluid.js:215-218
  encodeRadiation(pass) { addon.radiation(this.h); }                 // js/fluid.js:220-222
  encodeStep(pass) { addon.step(this.h, 1); }                        // js/fluid.js:225-253
  encodeSteps(pass, n) { addon.step(this.h, n); }

  // js/fluid.js:256-292: returns the stage names; the per-stage times (ms,
  // summed over the steps) are left in this.lastProfile
  encodeStepProfiled(encoder, querySet, base) {
    const ms = addon.profileSteps(this.h, 1);
    this.lastProfile = {};
    STAGES.forEach((s, i) => { this.lastProfile[s] = ms[i]; });
    return STAGES.slice();
  }

  // js/fluid.js:315-324: copy acc/rad/faceW to a free ring slot and clear acc;
  // a busy slot skips the readback
  encodeReadback(encoder) { this.pendingStage = addon.encodeReadback(this.h); }

  // js/fluid.js:327-342: cb({acc, rad, faceW}, tag) once the copy has landed
  afterSubmit(cb, tag) {
    const i = this.pendingStage;
    if (i < 0) return;
    this.pendingStage = -1;
    const h = this.h;
    const poll = () => {
      let r;
      try { r = addon.takeReadback(h, i, false); } catch (e) { return; }  // rejection swallowed as in the reference
      if (r === null) { setTimeout(poll, 1); return; }
      cb(r, tag);
    };
    setImmediate(poll);
  }

  // js/fluid.js:345-355: the raw buffer as a Float32Array
  async readBuffer(name) {
    if (name && typeof name === 'object') name = name.name;
    return new Float32Array(addon.readField(this.h, name));
  }

  writeBuffer(name, typedArray) { addon.writeField(this.h, name, typedArray); }
  synchronize() { addon.synchronize(this.h); }
  // throws when a NaN/Inf reached the fields or a residual target was missed
  status() { addon.status(this.h); }
}
