// Two frames of a reduced 2-D Karman scene (64x32, cylinder r = 4 at
// (16, 16)) through createGridSolver2 + createGridAdaptiveTimeStep2 in the
// Node binding, with the reference example's polygon helpers
// (examples/16-karman-vortex-street/main.js:217-240); tests/test_node.py
// compares the velocity with the oracle.  Usage: node karman_step.mjs <outdir>
import { writeFileSync, mkdirSync } from 'fs';
import * as flux from '../../tsl_array_n_amd/node/index.mjs';

const NX = 64, NY = 32, MARGIN = 1000;

function circle( cx, cy, r, segments ) {

	const verts = [];
	for ( let i = 0; i < segments; i ++ ) {

		const angle = ( i / segments ) * Math.PI * 2;
		verts.push( [ cx + r * Math.cos( angle ), cy + r * Math.sin( angle ) ] );

	}

	return verts;

}

function strip( innerX, outerX ) {

	return [ [ innerX, - MARGIN ], [ outerX, - MARGIN ], [ outerX, NY + MARGIN ], [ innerX, NY + MARGIN ] ];

}

async function main() {

	const dir = process.argv[ 2 ];
	mkdirSync( dir, { recursive: true } );
	await flux.init();
	const velocityGrid = flux.createFaceCenteredGrid2( NX, NY, 1, 1, 0, 0 );
	const random = flux.createJitterRandom( flux.DEFAULT_JITTER_SEED );
	const u0 = new Float32Array( ( NX + 1 ) * NY );
	for ( let f = 0; f < u0.length; f ++ ) u0[ f ] = 1 + ( random() * 2 - 1 ) * 1e-3;
	velocityGrid.dataU.fromArray( u0 );
	const collider = flux.createSDFStaticCollider2( NX, NY, 1, 1, 0, 0 );
	collider.addPolygon( circle( 16, 16, 4, 48 ) );
	const inflow = flux.createSDFInflow2( NX, NY, 1, 1, 0, 0, { velocity: [ 1, 0 ], mode: 'set' } );
	inflow.addPolygon( strip( 2, - MARGIN ) );
	const outflow = flux.createSDFOutflow2( NX, NY, 1, 1, 0, 0 );
	outflow.addPolygon( strip( NX - 2, NX + MARGIN ) );
	const dt = flux.array0( 'float' );
	dt.fromArray( new Float32Array( [ 1 ] ) );
	const solver = flux.createGridSolver2( {
		velocityGrid, gridSpacing: [ 1, 1 ], origin: [ 0, 0 ], collider, inflows: inflow, outflows: outflow,
		closedDomainBoundaryFlag: flux.DIRECTION_ALL & ~ flux.DIRECTION_RIGHT & ~ flux.DIRECTION_LEFT, dt,
		advection: { collider }, pressure: { multigrid: { numberOfLevels: 3 }, tolerance: 1e-6 }
	} );
	const stepper = flux.createGridAdaptiveTimeStep2( { velocityGrid: solver.velocityGrid, gridSpacing: [ 1, 1 ], dt,
		targetDt: 1, courantNumber: 0.5 } );
	const frames = [];
	for ( let f = 0; f < 2; f ++ ) {

		const k = await stepper.update();
		const diags = [];
		for ( let s = 0; s < k; s ++ ) diags.push( Object.assign( {}, await solver.onAdvanceTimeStep() ) );
		frames.push( { substeps: k, subDt: stepper.state.lastSubDt, diags } );

	}

	const save = async ( name, f ) => writeFileSync( `${ dir }/${ name }.bin`, Buffer.from( ( await f.toArray() ).buffer ) );
	writeFileSync( `${ dir }/u0.bin`, Buffer.from( u0.buffer ) );
	await save( 'u', solver.velocityGrid.dataU );
	await save( 'v', solver.velocityGrid.dataV );
	await save( 'collider', collider.grid.data );
	await save( 'inflow', inflow.grid.data );
	await save( 'outflow', outflow.grid.data );
	writeFileSync( `${ dir }/meta.json`, JSON.stringify( { frames } ) );
	console.log( 'node karman:', JSON.stringify( frames.map( ( f ) => [ f.substeps, f.diags.map( ( d ) => d.iterations ) ] ) ) );

}

main().catch( ( e ) => {

	console.error( e );
	process.exit( 1 );

} );
