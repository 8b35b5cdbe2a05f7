// Headless driver of the reference's CPU-side models (js/rigid.js, js/wood.js,
// js/contact.js), used only to write test fixtures in a container that holds
// the reference: it imports those modules from the reference checkout given
// on the command line and never runs on the GPU machine.
//
// It talks to a parent process over stdin/stdout, one JSON object per line:
//   {"cmd": "scene"}                      -> the default scene's pieces
//   {"cmd": "frame", "simDt": s, "readback": {acc, rad, faceW} | null, "rbDt": s}
//                                         -> the pieces of the next frame (simDt: this
//                                            frame's gas time; rbDt: that of the frame
//                                            the readback comes from)
// Each reply: {"pieces": [1280 floats], "n": highest used slot + 1,
//              "alive": number of pieces, "hrr": heat release estimate}.
//
// What it restates from the app (js/main.js), in its own code:
//  - the pre-lit default scene: a bed of seven coals, two logs and seven
//    sticks with preheated wood states (js/main.js:132-155);
//  - the frame's CPU work before the gas steps: rigid bodies, contact
//    conduction, the wood model fed by the last readback (js/main.js:261-264,
//    210-243);
//  - the piece record exactly as uploadPieces packs it (js/main.js:190-208).
import { pathToFileURL } from 'url';
import readline from 'readline';

const REF = process.argv[2];
const MAX_PIECES = 64, PIECE_FLOATS = 20, ACC_GLOBAL = 512;
const E_SCALE = 1e6, M_SCALE = 1e12, G_SCALE = 1e3;

async function main() {
  const load = p => import(pathToFileURL(REF + '/js/' + p).href);
  const { RigidWorld, Body } = await load('rigid.js');
  const { WoodPiece } = await load('wood.js');
  const { exchangeContacts } = await load('contact.js');
  const { PHYS, SCENE_WIDTH, CONTROL_DEFAULTS } = await load('config.js');

  const ctl = Object.assign({}, CONTROL_DEFAULTS);
  const world = new RigidWorld({ width: SCENE_WIDTH });
  world.p.muStatic = ctl.friction;
  world.p.muKinetic = 0.75 * ctl.friction;
  const bySlot = new Array(MAX_PIECES).fill(null);
  let live = [];
  let owners = [];

  // one wood piece in the first free slot (js/main.js:103-121)
  function place(x, y, side, heat) {
    const slot = bySlot.indexOf(null);
    if (slot < 0) return null;
    const wood = new WoodPiece(side, ctl.woodDensity);
    if (heat) wood.preheat(heat[0], heat[1], heat[2], ctl.charYield);
    const body = new Body(x, y, wood.half, (wood.mv + wood.mc) / (wood.side * wood.side), 0);
    world.add(body);
    const rec = { body, wood, slot, qConv: [0, 0, 0, 0], qInc: [1, 1, 1, 1].map(() => PHYS.sigma * PHYS.Tamb ** 4),
                  charPending: [0, 0, 0, 0] };
    body.userData = rec;
    bySlot[slot] = rec;
    live.push(rec);
    return rec;
  }

  // the pre-lit fire of js/main.js:132-155
  function defaultScene() {
    const cx = SCENE_WIDTH / 2, S = 0.0125;
    let x = cx - 0.06;
    for (const side of [0.016, 0.02, 0.015, 0.018, 0.02, 0.016, 0.017]) {
      const T = 1050;
      const r = place(x + side / 2, side / 2, side, [T, side, T - 30]);
      const w = r.wood;                         // an all-char coal
      w.mv = 0;
      w.mc = w.rhoC * side * side;
      w.updateSide();
      w.d = [side / 2, side / 2, side / 2, side / 2];
      w.charCov = [1, 1, 1, 1];
      w.Tp = [T - 20, T - 20, T - 20, T - 20];
      r.body.density = w.mc / (side * side);
      world.resize(r.body, w.half);
      x += side + 0.0015;
    }
    const lit = (px, py, side) => {
      const r = place(px, py, side, [880, side > 0.03 ? 0.003 : 0.002, 380]);
      r.wood.Tp = [590, 590, 590, 590];
    };
    lit(cx - 0.093, 0.025, 0.05);
    lit(cx + 0.093, 0.025, 0.05);
    const y1 = 0.02 + S + 0.001;
    for (const dx of [-0.045, -0.016, 0.013, 0.042]) lit(cx + dx, y1, 0.025);
    for (const dx of [-0.03, 0.0, 0.028]) lit(cx + dx, y1 + 2 * S + 0.002, 0.025);
  }

  // the piece record of js/main.js:190-208
  function pack() {
    const out = new Float32Array(MAX_PIECES * PIECE_FLOATS);
    let n = 0;
    for (const r of live) {
      const b = r.body, w = r.wood, o = r.slot * PIECE_FLOATS;
      out.set([b.x, b.y, b.a, b.half, b.vx, b.vy, b.w, 1], o);
      out.set(w.Ts, o + 8);
      out.set(w.mflux, o + 12);
      out.set(w.charCov, o + 16);
      n = Math.max(n, r.slot + 1);
    }
    owners = bySlot.slice();
    return { pieces: Array.from(out), n, alive: live.length };
  }

  // the readback into the wood model's inputs (js/main.js:210-230)
  function takeReadback(rb, simDt) {
    if (!rb || simDt <= 0) return 0;
    for (let s = 0; s < owners.length; s++) {
      const r = owners[s];
      if (!r || bySlot[s] !== r) continue;
      for (let f = 0; f < 4; f++) {
        const t = s * 4 + f;
        r.wood.exposure[f] = Math.min(1, rb.faceW[t] / 1e8 / Math.max(r.wood.side, 1e-4));
        r.qConv[f] = rb.acc[2 * t] / E_SCALE / simDt;
        r.charPending[f] += rb.acc[2 * t + 1] / M_SCALE;
        r.qInc[f] = rb.rad[4 * t];
      }
    }
    return rb.acc[ACC_GLOBAL] / G_SCALE / simDt;
  }

  // the wood step of every piece (js/main.js:232-243)
  function stepWood(dt) {
    if (dt <= 0) return;
    for (const r of live.slice()) {
      const w = r.wood;
      w.step(dt, r.qConv, r.qInc, r.charPending, ctl);
      r.charPending = [0, 0, 0, 0];
      if (w.dead) {
        world.remove(r.body);
        bySlot[r.slot] = null;
        live = live.filter(q => q !== r);
        continue;
      }
      r.body.density = (w.mv + w.mc) / (w.side * w.side);
      world.resize(r.body, w.half);
    }
  }

  const rl = readline.createInterface({ input: process.stdin });
  for await (const line of rl) {
    if (!line.trim()) continue;
    const m = JSON.parse(line);
    let reply;
    if (m.cmd === 'scene') {
      defaultScene();
      reply = pack();
    } else if (m.cmd === 'frame') {
      const hrr = takeReadback(m.readback, m.rbDt);
      world.advance(m.simDt);
      exchangeContacts(world, m.simDt);
      stepWood(m.simDt);
      reply = pack();
      reply.hrr = hrr;
    } else if (m.cmd === 'quit') {
      break;
    } else {
      reply = { error: 'unknown command ' + m.cmd };
    }
    process.stdout.write(JSON.stringify(reply) + '\n');
  }
}

main().catch(e => { console.error(e && e.stack || e); process.exit(1); });
