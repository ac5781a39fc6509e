// The 3-D collider exports through the Node binding: createSDFInflow3 in
// constrainVelocity, createOutflowPressureDirichlet3 + combineDirichlet3 in
// a pressure projection (evaluated on the GPU), and the advection solver
// with a collider; tests/test_node.py compares the outputs with the oracle.
// Usage: node collider3.mjs <outdir>
import { writeFileSync, mkdirSync } from 'fs';
import * as flux from '../../tsl_array_n_amd/node/index.mjs';

const save = async ( dir, name, field ) => writeFileSync( `${ dir }/${ name }.bin`, Buffer.from( ( await field.toArray() ).buffer ) );

async function main() {

	const dir = process.argv[ 2 ];
	mkdirSync( dir, { recursive: true } );
	await flux.init();
	const n = 12;
	const rnd = flux.createJitterRandom( 11 );
	const fill = ( f ) => {

		const a = new Float32Array( f.count );
		for ( let q = 0; q < a.length; q ++ ) a[ q ] = rnd() * 2 - 1;
		f.fromArray( a );

	};

	// --- inflow in constrainVelocity
	const vgrid = flux.createFaceCenteredGrid3( n, n, n, 1, 1, 1 );
	for ( const f of [ vgrid.dataU, vgrid.dataV, vgrid.dataW ] ) fill( f );
	for ( const c of [ 'U', 'V', 'W' ] ) await save( dir, 'vin_' + c, vgrid[ 'data' + c ] );
	const inflow = flux.createSDFInflow3( n, n, n, 1, 1, 1, 0, 0, 0, { velocity: [ 1.5, -0.25, 0.5 ], mode: 'set' } );
	inflow.addShapes( [ flux.sphere( [ 3.5, 6.0, 6.0 ], 2.5 ) ] );
	const bnd = flux.createGridBlockedBoundaryConditionSolver3( vgrid, n, n, n, 1, 1, 1, 0, 0, 0, null, [ inflow ] );
	bnd.constrainVelocity( 5 );
	await save( dir, 'inflow_sdf', inflow.grid.data );
	for ( const c of [ 'U', 'V', 'W' ] ) await save( dir, 'vout_' + c, vgrid[ 'data' + c ] );

	// --- outflow Dirichlet combined with a user pin
	const outflow = flux.createSDFOutflow3( n, n, n, 1, 1, 1 );
	outflow.addShapes( [ flux.box( [ 6.0, 11.0, 6.0 ], [ 3.0, 1.5, 3.0 ] ) ] );
	const pin = [ 2, 1, 3 ];
	const user = ( pos ) => ( { active: Math.floor( pos[ 0 ] ) === pin[ 0 ] && Math.floor( pos[ 1 ] ) === pin[ 1 ] && Math.floor( pos[ 2 ] ) === pin[ 2 ], target: 0.75 } );
	const dirichlet = flux.combineDirichlet3( flux.createOutflowPressureDirichlet3( [ outflow ] ), user );
	const solver = flux.createGridPressureSolver3( { resolution: [ n, n, n ], gridSpacing: [ 1, 1, 1 ], dirichlet,
		multigrid: { numberOfLevels: 2 }, tolerance: 1e-6 } );
	const pgrid = flux.createFaceCenteredGrid3( n, n, n, 1, 1, 1 );
	for ( const f of [ pgrid.dataU, pgrid.dataV, pgrid.dataW ] ) fill( f );
	const diag = Object.assign( {}, await solver.project( pgrid, pgrid )() );
	await save( dir, 'outflow_sdf', outflow.grid.data );
	await save( dir, 'mask', solver.dirichletMask );
	await save( dir, 'pressure', solver.pressure.data );
	// the host form of the same condition, at one cell inside the outflow and the pin
	const hostInside = dirichlet( [ 6.5, 11.5, 6.5 ] ), hostPin = dirichlet( [ 2.5, 1.5, 3.5 ] ), hostFree = dirichlet( [ 6.5, 5.5, 6.5 ] );

	// --- advection through a collider
	const collider = flux.createSDFStaticCollider3( n, n, n, 1, 1, 1 );
	collider.addShapes( [ flux.sphere( [ 6.0, 6.0, 6.0 ], 2.8 ) ] );
	const agrid = flux.createFaceCenteredGrid3( n, n, n, 1, 1, 1 );
	for ( const f of [ agrid.dataU, agrid.dataV, agrid.dataW ] ) fill( f );
	for ( const c of [ 'U', 'V', 'W' ] ) await save( dir, 'avel_' + c, agrid[ 'data' + c ] );
	const out = flux.createFaceCenteredGrid3( n, n, n, 1, 1, 1 );
	for ( const f of [ out.dataU, out.dataV, out.dataW ] ) f.fromArray( new Float32Array( f.count ).fill( -3 ) );
	const adv = flux.createSemiLagrangianAdvectionSolver3( { velocityGrid: agrid, collider, dt: 1.7, order: 2 } );
	adv.advectFaceCentered3( agrid, out )();
	await save( dir, 'collider_sdf', collider.grid.data );
	for ( const c of [ 'U', 'V', 'W' ] ) await save( dir, 'aout_' + c, out[ 'data' + c ] );
	writeFileSync( `${ dir }/meta.json`, JSON.stringify( { diag, pin, hostInside, hostPin, hostFree } ) );
	console.log( 'node collider3:', JSON.stringify( diag ) );

}

main().catch( ( e ) => {

	console.error( e );
	process.exit( 1 );

} );
