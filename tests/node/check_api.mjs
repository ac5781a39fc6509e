// CPU checks of the Node binding: exports, layout helpers, generator and
// seeding (compared by tests/test_node.py against the oracle), and the
// not-initialised error, a reference fixture read by loadSystem.
// Usage: node check_api.mjs <out.json> [fixture dir]
import { writeFileSync } from 'fs';
import * as flux from '../../tsl_array_n_amd/node/index.mjs';

const out = {};
out.exports = Object.keys( flux ).sort();
out.strides = flux.computeStrides( [ 4, 3, 2 ] );
try { flux.normalizeShape( [ 1, 2, 3, 4 ] ); out.shapeError = null; } catch ( e ) { out.shapeError = e.message; }
const rnd = flux.createJitterRandom( flux.DEFAULT_JITTER_SEED );
out.random = Array.from( { length: 16 }, () => rnd() );
const seed = flux.computeFlipBoxSeed3( { boxMin: [ 0, 0, 0 ], boxMax: [ 0.4, 0.6, 1.0 ], gridSpacingX: 1 / 16, gridSpacingY: 1 / 16, gridSpacingZ: 1 / 16 } );
out.seedCount = seed.count;
out.seedHead = Array.from( seed.positionsArray.subarray( 0, 24 ) );
let sum = 0;
for ( const v of seed.positionsArray ) sum += v;
out.seedSum = sum;
try { flux.arrayN( 'float', [ 4 ] ); out.initError = null; } catch ( e ) { out.initError = e.message; }
if ( process.argv[ 3 ] ) {

	const sys = flux.loadSystem( process.argv[ 3 ] );
	let bb = 0;
	for ( const v of sys.arrays.b ) bb += v * v;
	out.fixture = { resolution: sys.resolution, bLength: sys.arrays.b.length, weightU: sys.arrays.weightU.length,
		bNorm: Math.sqrt( bb ), iterations: sys.meta.diagnostics.iterations };
	const c = flux.computeAdaptiveSubSteps( 100, 1, 1 / 30, { courantNumber: 1 } );
	out.substeps = [ c.numSubSteps, c.subDt ];

}
writeFileSync( process.argv[ 2 ], JSON.stringify( out ) );
