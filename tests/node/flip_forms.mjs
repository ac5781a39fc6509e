// The 32^3 dam break through the Node binding with the benchmark's
// projection form (FP16 V-cycle, brick smoother with 12 sweeps, the
// free-cell coarse correction set by setCoarseSpace) and a second dt set by
// setDt, two steps each; the outputs are compared with the same run driven
// from Python by tests/test_node.py.
// Usage: node flip_forms.mjs <outdir>
import { writeFileSync, mkdirSync } from 'fs';
import * as flux from '../../tsl_array_n_amd/node/index.mjs';

const save = async ( dir, name, field ) => writeFileSync( `${ dir }/${ name }.bin`, Buffer.from( ( await field.toArray() ).buffer ) );

async function main() {

	const dir = process.argv[ 2 ];
	mkdirSync( dir, { recursive: true } );
	await flux.init();
	const n = 32, h = 1 / 32;
	const grid = flux.createFaceCenteredGrid3( n, n, n, h, h, h );
	const pos = [];
	for ( let k = 0; k < n; k ++ ) for ( let j = 0; j < 19; j ++ ) for ( let i = 0; i < 13; i ++ )
		for ( let sk = 0; sk < 2; sk ++ ) for ( let sj = 0; sj < 2; sj ++ ) for ( let si = 0; si < 2; si ++ )
			pos.push( ( i + ( si + 0.5 ) / 2 ) * h, ( j + ( sj + 0.5 ) / 2 ) * h, ( k + ( sk + 0.5 ) / 2 ) * h );
	const count = pos.length / 3;
	const solver = flux.createGridFlipSolver3( { velocityGrid: grid, maxParticles: count, dt: 1 / 60,
		pressure: { multigrid: { numberOfLevels: 3, storage: 'f16', smoother: 'brick',
			numberOfSmoothingIterationsDown: 12, numberOfSmoothingIterationsUp: 12, numberOfFinalIterations: 12 },
		tolerance: 1e-6 } } );
	solver.setCoarseSpace( true );
	solver.positions.fromArray( new Float32Array( pos ) );
	const diags = [];
	for ( let s = 0; s < 2; s ++ ) diags.push( Object.assign( {}, await solver.onAdvanceTimeStep() ) );
	solver.setDt( 1 / 120 );
	for ( let s = 0; s < 2; s ++ ) diags.push( Object.assign( {}, await solver.onAdvanceTimeStep() ) );
	await save( dir, 'positions', solver.positions );
	await save( dir, 'velocities', solver.velocities );
	writeFileSync( `${ dir }/diag.json`, JSON.stringify( { diags, count } ) );
	console.log( 'node flip forms:', JSON.stringify( diags ) );

}

main().catch( ( e ) => {

	console.error( e );
	process.exit( 1 );

} );
