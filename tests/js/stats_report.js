'use strict';
// Reads {forces, u0, area, settle, capture: [start, stop] | null, sim, cond,
// generated} as JSON from argv[2], folds the forces into the drop-in's
// ForceStats step by step and prints its CSV run report.
const fs = require('fs');
const path = require('path');
const { ForceStats } = require(path.join(__dirname, '..', '..', 'windtunnel_amd', 'js', 'stats.js'));
const inp = JSON.parse(fs.readFileSync(process.argv[2], 'utf8'));
const st = new ForceStats(inp.settle);
for (let k = 0; k < inp.forces.length; k++) {
  if (inp.capture && k === inp.capture[0]) st.setCapture(true, k);
  st.accumulate(inp.forces[k], inp.u0, inp.area, k + 1);
  if (inp.capture && k + 1 === inp.capture[1]) st.setCapture(false, k + 1);
}
const sim = Object.assign({}, inp.sim);
process.stdout.write(st.runCSV(sim, inp.cond, inp.generated));
