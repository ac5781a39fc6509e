// Three createGridSmokeSolver3 steps of a 24^3 plume through the Node
// binding; tests/test_node.py compares the fields with the oracle.
// Usage: node smoke_step.mjs <outdir>
import { writeFileSync, mkdirSync } from 'fs';
import * as flux from '../../tsl_array_n_amd/node/index.mjs';

async function main() {

	const dir = process.argv[ 2 ];
	mkdirSync( dir, { recursive: true } );
	await flux.init();
	const velocityGrid = flux.createFaceCenteredGrid3( 24, 24, 24, 1, 1, 1 );
	const smoke = flux.createGridSmokeSolver3( { velocityGrid, dt: 0.25, buoyancySmokeDensityFactor: - 0.05,
		buoyancyTemperatureFactor: 0.1, pressure: { multigrid: { numberOfLevels: 3 } } } );
	const diags = [];
	for ( let s = 0; s < 3; s ++ ) {

		smoke.sourceSphere( [ 12, 2.5, 12 ], 3.0, 1, 1 );
		diags.push( Object.assign( {}, await smoke.onAdvanceTimeStep() ) );

	}

	const save = async ( name, f ) => writeFileSync( `${ dir }/${ name }.bin`, Buffer.from( ( await f.toArray() ).buffer ) );
	await save( 'u', smoke.velocityGrid.dataU );
	await save( 'v', smoke.velocityGrid.dataV );
	await save( 'w', smoke.velocityGrid.dataW );
	await save( 'density', smoke.density.current );
	await save( 'temperature', smoke.temperature.current );
	writeFileSync( `${ dir }/meta.json`, JSON.stringify( { diags } ) );
	console.log( 'node smoke:', JSON.stringify( diags ) );

}

main().catch( ( e ) => {

	console.error( e );
	process.exit( 1 );

} );
