// A 3-D two-phase dam-break projection driven entirely through the Node
// binding, the way a fluxflow page would: seed, bin, beta, build u*, project.
// Usage: node projection.mjs <n> <outdir> [headline]
// "headline": the benchmark's form (FP16 V-cycle, brick smoother with 12
// sweeps, the free-cell coarse correction) instead of the reference's.
import { writeFileSync, mkdirSync } from 'fs';
import * as flux from '../../tsl_array_n_amd/node/index.mjs';

async function main() {

	const n = Number( process.argv[ 2 ] || 16 );
	const dir = process.argv[ 3 ];
	const headline = process.argv[ 4 ] === 'headline';
	mkdirSync( dir, { recursive: true } );
	await flux.init();
	const h = 1 / n;
	const grid = flux.createFaceCenteredGrid3( n, n, n, h, h, h );
	const seed = flux.computeFlipBoxSeed3( { boxMin: [ 0, 0, 0 ], boxMax: [ 0.4, 0.6, 1.0 ], gridSpacingX: h, gridSpacingY: h, gridSpacingZ: h } );
	const positions = flux.arrayN( 'vec3', seed.count ).fromArray( seed.positionsArray );
	const bins = flux.binParticles( grid, positions, seed.count );
	const fraction = flux.arrayN( 'float', [ n, n, n ] );
	const beta = flux.twoPhaseBeta( grid, bins, 1.0, 1000.0, fraction );
	const dt = 1 / 60;
	const v = new Float32Array( n * ( n + 1 ) * n );
	const gdt = Math.fround( Math.fround( -9.8 ) * Math.fround( dt ) );
	for ( let k = 0; k < n; k ++ ) for ( let j = 1; j < n; j ++ ) for ( let i = 0; i < n; i ++ ) v[ i + n * ( j + ( n + 1 ) * k ) ] = gdt;
	grid.dataV.fromArray( v );
	const pin = [ n >> 1, n - 1, n >> 1 ];
	const solver = flux.createGridPressureSolver3( {
		resolution: [ n, n, n ], gridSpacing: [ h, h, h ],
		dirichlet: ( pos ) => ( {
			active: Math.floor( pos[ 0 ] / h ) === pin[ 0 ] && Math.floor( pos[ 1 ] / h ) === pin[ 1 ] && Math.floor( pos[ 2 ] / h ) === pin[ 2 ],
			target: 0
		} ),
		faceWeights: beta, multigrid: headline ? {
			numberOfLevels: 3, storage: 'f16', smoother: 'brick', numberOfSmoothingIterationsDown: 12,
			numberOfSmoothingIterationsUp: 12, numberOfFinalIterations: 12
		} : { numberOfLevels: 3 }, tolerance: 1e-6, maxIterations: 100
	} );
	if ( headline ) solver.setCoarseSpace( true );
	const dispatch = solver.project( grid, grid );
	const diag = await dispatch();
	writeFileSync( `${ dir }/pressure.bin`, Buffer.from( ( await solver.pressure.data.toArray() ).buffer ) );
	writeFileSync( `${ dir }/u.bin`, Buffer.from( ( await grid.dataU.toArray() ).buffer ) );
	writeFileSync( `${ dir }/v.bin`, Buffer.from( ( await grid.dataV.toArray() ).buffer ) );
	writeFileSync( `${ dir }/w.bin`, Buffer.from( ( await grid.dataW.toArray() ).buffer ) );
	writeFileSync( `${ dir }/fraction.bin`, Buffer.from( ( await fraction.toArray() ).buffer ) );
	writeFileSync( `${ dir }/diag.json`, JSON.stringify( { diag, outside: bins.outside, count: seed.count } ) );
	console.log( 'node projection:', JSON.stringify( diag ) );

}

main().catch( ( e ) => {

	console.error( e );
	process.exit( 1 );

} );
