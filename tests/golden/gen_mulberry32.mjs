// Writes golden vectors of the reference's mulberry32 generator
// (packages/fluxflow/src/grid/jitter_random.js) for the oracle's tests.
// Run in a container that has the reference checkout:
//   node tests/golden/gen_mulberry32.mjs /root/reference > tests/golden/mulberry32.json
import { pathToFileURL } from 'url';
import path from 'path';

const root = process.argv[2] || '/root/reference';
import(pathToFileURL(path.join(root, 'packages/fluxflow/src/grid/jitter_random.js')).href).then((mod) => {
const seeds = [mod.DEFAULT_JITTER_SEED, 0, 1, 42, 0xFFFFFFFF];
const out = { defaultSeed: mod.DEFAULT_JITTER_SEED, vectors: [] };
for (const seed of seeds) {
	const random = mod.createJitterRandom(seed);
	const values = [];
	for (let i = 0; i < 64; i ++) values.push(random());
	out.vectors.push({ seed, values });
}
console.log(JSON.stringify(out));
});
