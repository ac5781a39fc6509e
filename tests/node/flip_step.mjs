// Two createGridFlipSolver3 steps of a 16^3 dam break through the Node
// binding, plus a collider set-up (SDF, face weights) and a constrained
// velocity; the outputs are compared with the oracle by tests/test_node.py.
// Usage: node flip_step.mjs <outdir>
import { writeFileSync, mkdirSync } from 'fs';
import * as flux from '../../tsl_array_n_amd/node/index.mjs';

const save = async ( dir, name, field ) => writeFileSync( `${ dir }/${ name }.bin`, Buffer.from( ( await field.toArray() ).buffer ) );

async function main() {

	const dir = process.argv[ 2 ];
	mkdirSync( dir, { recursive: true } );
	await flux.init();
	const n = 16, h = 1 / 16;
	const grid = flux.createFaceCenteredGrid3( n, n, n, h, h, h );
	const pos = [];
	for ( let k = 0; k < 16; k ++ ) for ( let j = 0; j < 10; j ++ ) for ( let i = 0; i < 6; i ++ )
		for ( let sk = 0; sk < 2; sk ++ ) for ( let sj = 0; sj < 2; sj ++ ) for ( let si = 0; si < 2; si ++ )
			pos.push( ( i + ( si + 0.5 ) / 2 ) * h, ( j + ( sj + 0.5 ) / 2 ) * h, ( k + ( sk + 0.5 ) / 2 ) * h );
	const count = pos.length / 3;
	const solver = flux.createGridFlipSolver3( { velocityGrid: grid, maxParticles: count, dt: 1 / 60,
		pressure: { multigrid: { numberOfLevels: 3 } } } );
	solver.positions.fromArray( new Float32Array( pos ) );
	const diags = [];
	for ( let s = 0; s < 2; s ++ ) diags.push( Object.assign( {}, await solver.onAdvanceTimeStep() ) );
	await save( dir, 'positions', solver.positions );
	await save( dir, 'velocities', solver.velocities );
	await save( dir, 'fluid', solver.fluidMask );

	// collider: a sphere on a 12^3 unit grid, its face weights, and one
	// constrained random velocity
	const collider = flux.createSDFStaticCollider3( 12, 12, 12, 1, 1, 1 );
	collider.addShapes( [ flux.sphere( [ 5.3, 6.1, 4.7 ], 2.6 ) ] );
	const weights = collider.computeFaceWeights( 12, 12, 12, 1, 1, 1 );
	const vgrid = flux.createFaceCenteredGrid3( 12, 12, 12, 1, 1, 1 );
	const rnd = flux.createJitterRandom( 7 );
	for ( const f of [ vgrid.dataU, vgrid.dataV, vgrid.dataW ] ) {

		const a = new Float32Array( f.count );
		for ( let q = 0; q < a.length; q ++ ) a[ q ] = rnd() * 2 - 1;
		f.fromArray( a );
		await save( dir, 'vin_' + f.shape.join( 'x' ), f );

	}

	const bnd = flux.createGridBlockedBoundaryConditionSolver3( vgrid, 12, 12, 12, 1, 1, 1, 0, 0, 0, collider, null, weights );
	bnd.closedDomainBoundaryFlag = 1 | 4 | 32;
	bnd.constrainVelocity( 5 );
	await save( dir, 'sdf', collider.grid.data );
	for ( const c of [ 'u', 'v', 'w' ] ) await save( dir, 'weight_' + c, weights[ c ] );
	await save( dir, 'vout_u', vgrid.dataU );
	await save( dir, 'vout_v', vgrid.dataV );
	await save( dir, 'vout_w', vgrid.dataW );
	writeFileSync( `${ dir }/meta.json`, JSON.stringify( { count, diags } ) );
	console.log( 'node flip:', JSON.stringify( diags ) );

}

main().catch( ( e ) => {

	console.error( e );
	process.exit( 1 );

} );
