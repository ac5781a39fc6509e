// The async boundary and the linalg exports through the Node binding:
//  * a 64^3 projection's Promise settles while the event loop keeps turning
//    (setImmediate callbacks run between the call and its settlement);
//  * createDotReducer's probes of examples/31-conjugate-gradient-3d
//    (single cells -> 1 in their own lane; a ramp -> 7140);
//  * createConjugateGradientSolver on A = diag(1..8), b = 1
//    (examples/04-conjugate-gradient).
// Usage: node async_linalg.mjs <out.json>
import { writeFileSync } from 'fs';
import * as flux from '../../tsl_array_n_amd/node/index.mjs';

async function main() {

	const out = process.argv[ 2 ];
	await flux.init();
	const res = {};

	// --- async projection
	const n = 64, h = 1 / n;
	const grid = flux.createFaceCenteredGrid3( n, n, n, h, h, h );
	const seed = flux.computeFlipBoxSeed3( { boxMin: [ 0, 0, 0 ], boxMax: [ 0.4, 0.6, 1.0 ], gridSpacingX: h, gridSpacingY: h, gridSpacingZ: h } );
	const positions = flux.arrayN( 'vec3', seed.count ).fromArray( seed.positionsArray );
	const bins = flux.binParticles( grid, positions, seed.count );
	const fraction = flux.arrayN( 'float', [ n, n, n ] );
	const beta = flux.twoPhaseBeta( grid, bins, 1.0, 1000.0, fraction );
	const v = new Float32Array( n * ( n + 1 ) * n );
	const gdt = Math.fround( Math.fround( -9.8 ) * Math.fround( 1 / 60 ) );
	for ( let k = 0; k < n; k ++ ) for ( let j = 1; j < n; j ++ ) for ( let i = 0; i < n; i ++ ) v[ i + n * ( j + ( n + 1 ) * k ) ] = gdt;
	grid.dataV.fromArray( v );
	const pin = [ n >> 1, n - 1, n >> 1 ];
	const solver = flux.createGridPressureSolver3( {
		resolution: [ n, n, n ], gridSpacing: [ h, h, h ],
		dirichlet: ( pos ) => ( {
			active: Math.floor( pos[ 0 ] / h ) === pin[ 0 ] && Math.floor( pos[ 1 ] / h ) === pin[ 1 ] && Math.floor( pos[ 2 ] / h ) === pin[ 2 ],
			target: 0
		} ),
		faceWeights: beta, multigrid: { numberOfLevels: 4 }, tolerance: 1e-6
	} );
	let ticks = 0, settled = false;
	const spin = () => {

		if ( settled ) return;
		ticks ++;
		setImmediate( spin );

	};

	setImmediate( spin );
	const t0 = Date.now();
	const diag = await solver.project( grid, grid )();
	settled = true;
	res.project = { diag, ticks, ms: Date.now() - t0 };

	// two projections queued at once are serialised, both settle
	const both = await Promise.all( [ solver.project( grid, grid )(), solver.project( grid, grid )() ] );
	res.both = both.map( ( d ) => d.converged );

	// the free-cell coarse-space correction cuts the iterations; the
	// projection is accepted at the same tolerance
	grid.dataV.fromArray( v );
	grid.dataU.fromArray( new Float32Array( ( n + 1 ) * n * n ) );
	grid.dataW.fromArray( new Float32Array( n * n * ( n + 1 ) ) );
	solver.pressure.data.fromArray( new Float32Array( n * n * n ) );
	const cold = await solver.project( grid, grid )();
	res.coldIterations = cold.iterations;
	solver.setCoarseSpace( true, [] );
	grid.dataV.fromArray( v );
	grid.dataU.fromArray( new Float32Array( ( n + 1 ) * n * n ) );
	grid.dataW.fromArray( new Float32Array( n * n * ( n + 1 ) ) );
	solver.pressure.data.fromArray( new Float32Array( n * n * n ) );
	const cs = await solver.project( grid, grid )();
	res.coarseSpace = { converged: cs.converged, iterations: cs.iterations };

	// --- dot reducer probes
	const shape = [ 6, 4, 5 ];
	const ones = flux.arrayN( 'float', shape ).fromArray( new Float32Array( 120 ).fill( 1 ) );
	const delta = flux.arrayN( 'float', shape );
	const reducer = flux.createDotReducer( shape, delta, ones );
	res.lanes = reducer.lanes;
	res.probes = [];
	for ( const [ pi, pj, pk ] of [ [ 0, 0, 0 ], [ 0, 0, 1 ], [ 0, 1, 0 ], [ 5, 3, 4 ], [ 2, 3, 1 ] ] ) {

		const d = new Float32Array( 120 );
		d[ pi + 6 * ( pj + 4 * pk ) ] = 1;
		delta.fromArray( d );
		const total = await reducer.read();
		const lanes = Array.from( await reducer.partial.toArray() );
		res.probes.push( { cell: [ pi, pj, pk ], total, lit: lanes.map( ( x, l ) => ( x !== 0 ? [ l, x ] : null ) ).filter( Boolean ) } );

	}

	const ramp = flux.arrayN( 'float', shape ).fromArray( Float32Array.from( { length: 120 }, ( _, i ) => i ) );
	res.ramp = await flux.createDotReducer( shape, ramp, ones ).read();

	// --- plain CG on a diagonal operator
	const N = 8;
	const b = flux.arrayN( 'float', N ).fromArray( new Float32Array( N ).fill( 1 ) );
	const x = flux.arrayN( 'float', N ).fromArray( new Float32Array( N ) );
	const d = flux.arrayN( 'float', N ).fromArray( Float32Array.from( { length: N }, ( _, i ) => i + 1 ) );
	const cg = flux.createConjugateGradientSolver( { diagonal: d }, b, x );
	res.cgSucceeded = await cg.solve( 1e-5, 20 );
	res.cgX = Array.from( await x.toArray() );
	writeFileSync( out, JSON.stringify( res ) );
	console.log( JSON.stringify( res.project ) );

}

main().catch( ( e ) => {

	console.error( e );
	process.exit( 1 );

} );
