'use strict';
// Runs the JavaScript CPU path (splat_oracle.js) over files, for the parity tests and the CPU
// baseline. TEST INFRASTRUCTURE ONLY.
//
//   node oracle/run_js.js RECORDS.bin VIEW.bin OUT_PREFIX        -> OUT_keys.bin (i32), OUT_order.bin (u32),
//                                                                   OUT_cov.bin (f64 n x 3)
//   node oracle/run_js.js RECORDS.bin VIEW.bin --time SECONDS     -> one JSON line: splats/s, median of runs
//
// VIEW.bin holds the 37 floats of the per-view uniform (the wd_view layout).
const fs = require('fs');
const os = require('os');
const oracle = require('./splat_oracle.js');

function main(argv) {
  const bytes = new Uint8Array(fs.readFileSync(argv[0]));
  const vb = fs.readFileSync(argv[1]);
  const view = oracle.viewFromFloats(new Float32Array(vb.buffer.slice(vb.byteOffset, vb.byteOffset + 148)));
  const n = bytes.length / 32;
  if (argv[2] === '--time') {
    const budget = Number(argv[3] || 10);
    const times = [];
    oracle.run(bytes, view);  // warm the JIT
    const start = process.hrtime.bigint();
    let survivors = 0;
    do {
      const t0 = process.hrtime.bigint();
      survivors = oracle.run(bytes, view).length;
      times.push(Number(process.hrtime.bigint() - t0) / 1e9);
    } while (times.length < 5 && Number(process.hrtime.bigint() - start) / 1e9 < budget);
    times.sort((a, b) => a - b);
    const median = times[times.length >> 1];
    const cpus = os.cpus();
    console.log(JSON.stringify({
      splats: n, survivors, runs: times.length, median_s: median, splats_per_s: n / median,
      threads: 1, host_cpus: cpus.length, cpu_model: cpus.length ? cpus[0].model : '', node: process.version,
    }));
    return;
  }
  const cov = new Float64Array(n * 3).fill(NaN);
  const k = oracle.keys(bytes, view, cov);
  const order = oracle.sort(k);
  const out = argv[2];
  fs.writeFileSync(out + '_keys.bin', Buffer.from(k.buffer));
  fs.writeFileSync(out + '_order.bin', Buffer.from(order.buffer));
  fs.writeFileSync(out + '_cov.bin', Buffer.from(cov.buffer));
}

main(process.argv.slice(2));
