// createSemiLagrangianAdvectionSolver3 through the Node binding: a face grid
// and a cell-centred scalar advected by MacCormack (order 2) through a
// separate velocity grid, with dt as a live array0 field; tests/test_node.py
// compares the results with the oracle.  Usage: node advect_step.mjs <outdir>
import { writeFileSync, mkdirSync } from 'fs';
import * as flux from '../../tsl_array_n_amd/node/index.mjs';

const N = [ 20, 16, 12 ];

function fill( shape, f ) {

	const out = new Float32Array( shape[ 0 ] * shape[ 1 ] * shape[ 2 ] );
	for ( let k = 0; k < shape[ 2 ]; k ++ ) for ( let j = 0; j < shape[ 1 ]; j ++ ) for ( let i = 0; i < shape[ 0 ]; i ++ ) {

		out[ i + shape[ 0 ] * ( j + shape[ 1 ] * k ) ] = f( i, j, k );

	}

	return out;

}

async function main() {

	const dir = process.argv[ 2 ];
	mkdirSync( dir, { recursive: true } );
	await flux.init();
	const vel = flux.createFaceCenteredGrid3( N[ 0 ], N[ 1 ], N[ 2 ], 1, 1, 1 );
	const data = flux.createFaceCenteredGrid3( N[ 0 ], N[ 1 ], N[ 2 ], 1, 1, 1 );
	const out = flux.createFaceCenteredGrid3( N[ 0 ], N[ 1 ], N[ 2 ], 1, 1, 1 );
	const arrays = {};
	const comps = [ 'U', 'V', 'W' ];
	comps.forEach( ( c, a ) => {

		const shape = vel[ 'dataSize' + c ];
		arrays[ 'vel' + c ] = fill( shape, ( i, j, k ) => 1.1 * Math.sin( 0.3 * i + a ) * Math.cos( 0.2 * j ) * Math.cos( 0.25 * k + 0.5 * a ) + 0.3 );
		arrays[ 'data' + c ] = fill( shape, ( i, j, k ) => Math.sin( 0.5 * i ) * Math.cos( 0.35 * j + 0.2 * k ) + ( i + 0.7 * j > 12 ? 2 : 0 ) + 0.1 * a );
		vel[ 'data' + c ].fromArray( arrays[ 'vel' + c ] );
		data[ 'data' + c ].fromArray( arrays[ 'data' + c ] );

	} );
	const dt = flux.array0( 'float' );
	dt.fromArray( new Float32Array( [ 1.3 ] ) );
	const solver = flux.createSemiLagrangianAdvectionSolver3( { velocityGrid: vel, dt, order: 2 } );
	solver.advectFaceCentered3( data, out )();
	const sIn = flux.createCellCenteredScalarGrid3( N[ 0 ], N[ 1 ], N[ 2 ], 1, 1, 1 );
	const sOut = flux.createCellCenteredScalarGrid3( N[ 0 ], N[ 1 ], N[ 2 ], 1, 1, 1 );
	arrays.scalar = fill( N, ( i, j, k ) => ( ( i * 7 + j * 3 + k * 5 ) % 11 ) / 11 );
	sIn.data.fromArray( arrays.scalar );
	dt.fromArray( new Float32Array( [ 2.2 ] ) );  // read at dispatch
	solver.advectScalar3( sIn, sOut )();
	for ( const name of Object.keys( arrays ) ) writeFileSync( `${ dir }/${ name }.bin`, Buffer.from( arrays[ name ].buffer ) );
	for ( const c of comps ) writeFileSync( `${ dir }/out${ c }.bin`, Buffer.from( ( await out[ 'data' + c ].toArray() ).buffer ) );
	writeFileSync( `${ dir }/outScalar.bin`, Buffer.from( ( await sOut.data.toArray() ).buffer ) );
	console.log( 'node advect: done' );

}

main().catch( ( e ) => {

	console.error( e );
	process.exit( 1 );

} );
