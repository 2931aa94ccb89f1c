// Writes mulberry32_vectors.json: the first draws of Mulberry32 (Tommy Ettinger's public
// generator, in its published JavaScript form with Math.imul and unsigned shifts) for a
// handful of seeds. An implementation independent of the oracle's C u32 arithmetic, run
// with node: `node tests/golden/make_mulberry32_vectors.js > tests/golden/mulberry32_vectors.json`.
function mulberry32(a) {
  return function () {
    a = (a + 0x6D2B79F5) | 0;
    let t = Math.imul(a ^ (a >>> 15), 1 | a);
    t = (t + Math.imul(t ^ (t >>> 7), 61 | t)) ^ t;
    return ((t ^ (t >>> 14)) >>> 0) / 4294967296;
  };
}
const seeds = [0, 1, 7, 0x051042, 123456789, 0x7fffffff, 0x80000000, 0xffffffff];
const out = {};
for (const s of seeds) {
  const next = mulberry32(s >>> 0);
  const draws = [];
  for (let k = 0; k < 64; k++) draws.push(next());
  out[String(s >>> 0)] = draws;
}
// doubles of the form u32 / 2^32 print exactly with 17 significant digits
process.stdout.write(JSON.stringify(out, (k, v) => (typeof v === "number" ? Number(v.toPrecision(17)) : v)) + "\n");
