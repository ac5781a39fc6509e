// Two createGridFlipSolver3 steps with carried concentration, density
// coupling and the top-centre pressure pin through the Node binding, from
// positions and concentrations the test writes to <dir>/pos.bin and
// <dir>/conc.bin; tests/test_node.py compares the results with the oracle.
// Usage: node flip_conc_step.mjs <dir> <nx> <ny> <nz>
import { readFileSync, writeFileSync } from 'fs';
import * as flux from '../../tsl_array_n_amd/node/index.mjs';

async function main() {

	const [ dir, nx, ny, nz ] = [ process.argv[ 2 ], + process.argv[ 3 ], + process.argv[ 4 ], + process.argv[ 5 ] ];
	await flux.init();
	const pos = new Float32Array( readFileSync( `${ dir }/pos.bin` ).buffer.slice( 0 ) );
	const conc = new Float32Array( readFileSync( `${ dir }/conc.bin` ).buffer.slice( 0 ) );
	const velocityGrid = flux.createFaceCenteredGrid3( nx, ny, nz, 1, 1, 1 );
	const solver = flux.createGridFlipSolver3( {
		velocityGrid, maxParticles: conc.length, dt: 0.05, carryConcentration: true, mixing: 0.3, fade: 0.01,
		ambientDensity: 1, componentDensity: 3, pressurePin: 'topCenter', resample: { enabled: false },
		pressure: { multigrid: { numberOfLevels: 2 } }
	} );
	solver.positions.fromArray( pos );
	solver.concentration.fromArray( conc );
	const diags = [];
	for ( let s = 0; s < 2; s ++ ) diags.push( Object.assign( {}, await solver.onAdvanceTimeStep() ) );
	const save = async ( name, f ) => writeFileSync( `${ dir }/${ name }.bin`, Buffer.from( ( await f.toArray() ).buffer ) );
	await save( 'out_pos', solver.positions );
	await save( 'out_vel', solver.velocities );
	await save( 'out_conc', solver.concentration );
	await save( 'out_cell_density', solver.cellDensity );
	writeFileSync( `${ dir }/meta.json`, JSON.stringify( { diags } ) );
	console.log( 'node flip conc:', JSON.stringify( diags ) );

}

main().catch( ( e ) => {

	console.error( e );
	process.exit( 1 );

} );
