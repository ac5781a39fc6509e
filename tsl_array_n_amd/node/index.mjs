// tsl_array_n / fluxflow API on the MI355X library (ESM, Node >= 12).
//
// A drop-in for the reference's projection path: the same factory names,
// argument meanings and error messages as
//   packages/tsl_array_n/src/{context,array}.js            (init, arrayN, array0)
//   packages/fluxflow/src/linalg/{multigrid,linalg}.js      (operator, preconditioners, PCG)
//   packages/fluxflow/src/grid/grid_data3.js                (createFaceCenteredGrid3)
//   packages/fluxflow/src/grid/grid_pressure_solver3.js     (createGridPressureSolver3)
//   packages/fluxflow/src/grid/grid_flip_solver3.js:70-130  (computeFlipBoxSeed3)
//   packages/fluxflow/src/grid/jitter_random.js             (createJitterRandom)
// What the reference expresses as shader functions (a Dirichlet mask, face
// weights, `dirichlet(pos)`) is passed here as fields, or as a JavaScript
// function evaluated per cell on the host.
import { createRequire } from 'module';

const require = createRequire( import.meta.url );
const addon = require( '../lib/fluxflow_napi.node' );

const TYPE_CODES = { float: 0, int: 1, u8: 2, f64: 3, vec3: 6 };
const TYPED = { float: Float32Array, int: Int32Array, u8: Uint8Array, f64: Float64Array, vec3: Float32Array };
const PRECONDITIONERS = { multigrid: 0, jacobi: 1, none: 2 };

export const DEFAULT_JITTER_SEED = 0x9E3779B9;

// ------------------------------------------------------------- context
export async function init( options = {} ) {

	addon.init( options.device === undefined ? 0 : options.device );
	return { backend: 'hip', device: addon.deviceInfo() };

}

export function sync() {

	addon.sync();

}

// -------------------------------------------------------------- fields
export function normalizeShape( shape ) {

	const dims = typeof shape === 'number' ? [ shape ] : Array.from( shape );
	if ( dims.length < 1 || dims.length > 3 ) throw new Error( `arrayN: only 1D/2D/3D shapes are supported, got ${ dims.length }D.` );
	for ( const d of dims ) {

		if ( ! Number.isInteger( d ) || d <= 0 ) throw new Error( `arrayN: shape entries must be positive integers, got ${ d }.` );

	}

	return dims;

}

// first index fastest: strides [1, n0, n0*n1] (array.js:24-38)
export function computeStrides( shape ) {

	const strides = [];
	let s = 1;
	for ( const d of shape ) {

		strides.push( s );
		s *= d;

	}

	return strides;

}

function wrapField( handle, type, shape ) {

	const strides = computeStrides( shape );
	const count = shape.reduce( ( a, n ) => a * n, 1 );
	// callable like the reference's field accessor: here it yields the flat index
	const field = ( ...I ) => I.reduce( ( acc, v, a ) => acc + v * strides[ a ], 0 );
	field.handle = handle;
	field.type = type;
	field.shape = shape;
	field.count = count;
	field.strides = strides;
	field.at = field;
	field.toArray = async () => {

		const out = new TYPED[ type ]( type === 'vec3' ? count * 3 : count );
		addon.fieldDownload( handle, out );
		return out;

	};

	field.fromArray = ( typed ) => {

		addon.fieldUpload( handle, typed );
		return field;

	};

	field.fill = ( value ) => {

		addon.fieldFill( handle, value );
		return field;

	};

	return field;

}

export function arrayN( type, shape ) {

	if ( ! ( type in TYPE_CODES ) ) throw new Error( `arrayN: unknown type "${ type }".` );
	const dims = normalizeShape( shape );
	return wrapField( addon.fieldCreate( TYPE_CODES[ type ], dims ), type, dims );

}

export function array0( type ) {

	return arrayN( type, [ 1 ] );

}

// ---------------------------------------------------------------- linalg
function handles( fields ) {

	return fields ? fields.map( ( f ) => f.handle ) : null;

}

export function createLaplacianOperator( shape, gridSpacing, options = {} ) {

	const op = addon.operatorCreate( shape, gridSpacing, options.dirichletMask ? options.dirichletMask.handle : null,
		handles( options.faceWeights ) );
	function applyLaplacian( input, output ) {

		return function dispatch() {

			addon.operatorApply( op, input.handle, output.handle );

		};

	}

	applyLaplacian.native = { kind: 'operator', op };
	return applyLaplacian;

}

export function createMultigridPreconditioner( shape, gridSpacing, options = {} ) {

	if ( shape.length < 1 || shape.length > 3 ) throw new Error( `multigrid: only 1D/2D/3D shapes are supported, got ${ shape.length }D.` );
	if ( gridSpacing.length !== shape.length ) throw new Error( `multigrid: gridSpacing length (${ gridSpacing.length }) must match shape length (${ shape.length }).` );
	const mg = addon.multigridCreate( shape, gridSpacing, options,
		options.dirichletMask ? options.dirichletMask.handle : null,
		options.coarseDirichletMask ? options.coarseDirichletMask.handle : null,
		handles( options.faceWeights ) );
	function applyMultigridPreconditioner( input, output ) {

		return function dispatch() {

			addon.multigridApply( mg, input.handle, output.handle );

		};

	}

	applyMultigridPreconditioner.native = { kind: 'multigrid', mg };
	applyMultigridPreconditioner.refreshCoarseLevels = () => addon.multigridRefresh( mg );
	applyMultigridPreconditioner.options = Object.assign( {
		numberOfLevels: 1, numberOfSmoothingIterationsDown: 2, numberOfSmoothingIterationsUp: 2,
		numberOfCoarsestIterations: 20, numberOfFinalIterations: 2, sorFactor: 1.0
	}, options );
	return applyMultigridPreconditioner;

}

export function createJacobiPreconditioner() {

	const f = () => () => {

		throw new Error( 'createJacobiPreconditioner: only usable inside the conjugate-gradient solver here' );

	};

	f.native = { kind: 'jacobi' };
	return f;

}

export function createIdentityPreconditioner() {

	const f = () => () => {

		throw new Error( 'createIdentityPreconditioner: only usable inside the conjugate-gradient solver here' );

	};

	f.native = { kind: 'none' };
	return f;

}

export function createPreconditionedConjugateGradientSolver( applyOperator, applyPreconditioner, b, x, options = {} ) {

	if ( b.type !== x.type ) throw new Error( `conjugateGradient: element type mismatch, b.type(${ b.type }) != x.type(${ x.type }).` );
	if ( b.shape.length !== x.shape.length || b.shape.some( ( v, i ) => v !== x.shape[ i ] ) ) {

		throw new Error( `conjugateGradient: shape mismatch, b.shape(${ b.shape }) != x.shape(${ x.shape }).` );

	}

	if ( b.type !== 'float' ) throw new Error( `conjugateGradient: GPU atomic dot product only supports type "float", got "${ b.type }".` );
	if ( ! applyOperator || ! applyOperator.native ) throw new Error( 'conjugateGradient: the operator must come from createLaplacianOperator' );
	const pre = applyPreconditioner && applyPreconditioner.native ? applyPreconditioner.native : { kind: 'none' };
	const kind = pre.kind === 'multigrid' ? PRECONDITIONERS.multigrid : pre.kind === 'jacobi' ? PRECONDITIONERS.jacobi : PRECONDITIONERS.none;
	const solver = addon.pcgCreate( applyOperator.native.op, kind, pre.kind === 'multigrid' ? pre.mg : null, b.handle, x.handle );
	const state = { residualSquared: 0, iterations: 0, stoppedBy: 'none' };
	const settings = { reportNoiseFloor: options.reportNoiseFloor === true };
	// solve(tol, maxiter, residualCheckInterval, gpuResidentScalars, batchIterations,
	//       gpuResidentSetup, relativeTolerance, ...) -- linalg.js:2176
	async function solve( tol, maxiter, residualCheckInterval = 1, gpuResidentScalars = false, batchIterations = true,
		gpuResidentSetup = false, relativeTolerance = false ) {

		const st = await addon.pcgSolveAsync( solver, tol, maxiter, relativeTolerance === true );
		state.residualSquared = st.residualSquared;
		state.iterations = st.iterations;
		state.stoppedBy = st.stoppedBy === 'converged' ? 'none' : st.stoppedBy;
		return st.converged;

	}

	return { solve, state, settings };

}

// createDotReducer(shape, fieldA, fieldB) -- linalg.js:374-512: read()
// resolves to sum(a * b); lanes and the per-lane partial field as in the
// reference; dispatch() computes the lanes without reading back.
export function createDotReducer( shape, fieldA, fieldB ) {

	if ( shape.length < 1 || shape.length > 3 ) throw new Error( `createDotReducer: only 1D/2D/3D shapes are supported, got ${ shape.length }D.` );
	const reducer = addon.dotCreate( shape, fieldA.handle, fieldB.handle );
	const info = addon.dotInfo( reducer );
	const partial = wrapField( info.partial, 'float', [ info.lanes ] );
	partial.keepAlive = reducer;
	return {
		read: () => addon.dotReadAsync( reducer ),
		lanes: info.lanes,
		partial,
		dispatch: () => addon.dotDispatch( reducer )
	};

}

// createConjugateGradientSolver(applyOperator, b, x) -- linalg.js:563-728.
// applyOperator is a createLaplacianOperator result, or { diagonal: field }
// for the diagonal operator out = diag * in (examples/04-conjugate-gradient).
export function createConjugateGradientSolver( applyOperator, b, x, options = {} ) {

	if ( b.type !== x.type ) throw new Error( `conjugateGradient: element type mismatch, b.type(${ b.type }) != x.type(${ x.type }).` );
	if ( b.shape.length !== x.shape.length || b.shape.some( ( v, i ) => v !== x.shape[ i ] ) ) {

		throw new Error( `conjugateGradient: shape mismatch, b.shape(${ b.shape }) != x.shape(${ x.shape }).` );

	}

	if ( b.type !== 'float' ) throw new Error( `conjugateGradient: GPU atomic dot product only supports type "float", got "${ b.type }".` );
	const op = applyOperator && applyOperator.native ? applyOperator.native.op : null;
	const diag = applyOperator && applyOperator.diagonal ? applyOperator.diagonal.handle : null;
	if ( ! op && ! diag ) throw new Error( 'conjugateGradient: the operator must come from createLaplacianOperator or be { diagonal: field }' );
	const solver = addon.cgCreate( op, diag, b.handle, x.handle );
	const f = addon.cgFields( solver );
	const field = ( h ) => {

		const w = wrapField( h, 'float', b.shape );
		w.keepAlive = solver;
		return w;

	};

	async function solve( tol, maxiter ) {

		const r = await addon.cgSolveAsync( solver, tol, maxiter );
		return r.converged;

	}

	return { solve, p: field( f.p ), r: field( f.r ), Ap: field( f.Ap ), Ax: field( f.Ax ) };

}

// ------------------------------------------------------------ grid data
export function createFaceCenteredGrid3( nx, ny, nz, hx, hy, hz, ox = 0, oy = 0, oz = 0 ) {

	const dataSizeU = [ nx + 1, ny, nz ];
	const dataSizeV = [ nx, ny + 1, nz ];
	const dataSizeW = [ nx, ny, nz + 1 ];
	const grid = {
		resolution: [ nx, ny, nz ], gridSpacing: [ hx, hy, hz ], origin: [ ox, oy, oz ],
		dataSizeU, dataSizeV, dataSizeW,
		dataOriginU: [ ox, oy + hy * 0.5, oz + hz * 0.5 ],
		dataOriginV: [ ox + hx * 0.5, oy, oz + hz * 0.5 ],
		dataOriginW: [ ox + hx * 0.5, oy + hy * 0.5, oz ],
		dataU: arrayN( 'float', dataSizeU ),
		dataV: arrayN( 'float', dataSizeV ),
		dataW: arrayN( 'float', dataSizeW )
	};
	grid.clear = () => {

		grid.dataU.fill( 0 );
		grid.dataV.fill( 0 );
		grid.dataW.fill( 0 );

	};

	return grid;

}

// ------------------------------------------------------- pressure solver
const DEFAULT_MAX_PLAUSIBLE_PRESSURE = 1e6;
const FIELD_PRESSURE = 0, FIELD_RHS = 1, FIELD_MASK = 2, FIELD_VENT = 3;

function evaluateDirichlet( dirichlet, resolution, gridSpacing, origin ) {

	// the reference evaluates dirichlet(pos) at cell centres (dataPosition,
	// grid_pressure_solver3.js:177-181); a JavaScript function is evaluated here
	const [ nx, ny, nz ] = resolution;
	const active = new Uint8Array( nx * ny * nz );
	const target = new Float32Array( nx * ny * nz );
	for ( let k = 0; k < nz; k ++ ) for ( let j = 0; j < ny; j ++ ) for ( let i = 0; i < nx; i ++ ) {

		const pos = [ origin[ 0 ] + ( i + 0.5 ) * gridSpacing[ 0 ], origin[ 1 ] + ( j + 0.5 ) * gridSpacing[ 1 ], origin[ 2 ] + ( k + 0.5 ) * gridSpacing[ 2 ] ];
		const r = dirichlet( pos );
		const c = i + nx * ( j + ny * k );
		active[ c ] = r.active ? 1 : 0;
		target[ c ] = r.target || 0;

	}

	return { active, target };

}

export function createGridPressureSolver3( {
	resolution, gridSpacing, origin = [ 0, 0, 0 ],
	dirichlet, faceWeights, colliderWeights,
	multigrid = {}, preconditioner = 'multigrid',
	tolerance = 1e-6, maxIterations = null, maxIterationsFactor = 1.5, relativeTolerance = true,
	maxPlausiblePressure = DEFAULT_MAX_PLAUSIBLE_PRESSURE
} = {} ) {

	if ( ! Object.prototype.hasOwnProperty.call( PRECONDITIONERS, preconditioner ) ) {

		throw new Error( `grid_pressure_solver3: unknown preconditioner '${ preconditioner }' (expected ${ Object.keys( PRECONDITIONERS ).join( ', ' ) }).` );

	}

	let dirichletActive = null, dirichletTarget = null;
	if ( typeof dirichlet === 'function' && dirichlet.evaluateFields ) {

		// outflow Dirichlet conditions evaluated on the GPU
		( { active: dirichletActive, target: dirichletTarget } = dirichlet.evaluateFields( resolution, gridSpacing, origin ) );

	} else if ( typeof dirichlet === 'function' ) {

		const { active, target } = evaluateDirichlet( dirichlet, resolution, gridSpacing, origin );
		dirichletActive = arrayN( 'u8', resolution ).fromArray( active );
		dirichletTarget = arrayN( 'float', resolution ).fromArray( target );

	} else if ( dirichlet ) {

		dirichletActive = dirichlet.active;
		dirichletTarget = dirichlet.target || null;

	}

	const triple = ( w ) => ( w ? [ w.u, w.v, w.w ].map( ( f ) => f.handle ) : null );
	const solver = addon.pressure3Create( {
		resolution, gridSpacing, origin,
		dirichletActive: dirichletActive ? dirichletActive.handle : null,
		dirichletTarget: dirichletTarget ? dirichletTarget.handle : null,
		faceWeights: triple( faceWeights ),
		colliderWeights: triple( colliderWeights ),
		multigrid, preconditionerKind: PRECONDITIONERS[ preconditioner ],
		tolerance, maxIterations, maxIterationsFactor, relativeTolerance, maxPlausiblePressure
	} );
	const view = ( which, type ) => {

		const h = addon.pressure3Field( solver, which );
		return h ? wrapField( h, type, resolution.slice() ) : null;

	};

	const pressure = { resolution, gridSpacing, origin, data: view( FIELD_PRESSURE, 'float' ) };
	const diagnostics = { converged: null, rejected: false, iterations: null, stoppedBy: null };
	const settings = { tolerance, preconditioner, multigrid };
	const keep = { dirichletActive, dirichletTarget, faceWeights, colliderWeights };

	function project( input, output ) {

		if ( input.resolution.join() !== resolution.join() ) {

			throw new Error( `createGridPressureSolver3: project() input resolution [${ input.resolution }] does not match constructed resolution [${ resolution }].` );

		}

		const inFields = [ input.dataU.handle, input.dataV.handle, input.dataW.handle ];
		const outFields = [ output.dataU.handle, output.dataV.handle, output.dataW.handle ];
		return async function dispatch() {

			const d = await addon.pressure3ProjectAsync( solver, inFields, outFields );
			diagnostics.converged = d.converged;
			diagnostics.rejected = d.rejected;
			diagnostics.iterations = d.iterations;
			diagnostics.residual = d.residual;
			diagnostics.stoppedBy = d.stoppedBy === 'converged' ? 'none' : d.stoppedBy;
			diagnostics.badCells = d.badCells;
			return diagnostics;

		};

	}

	// no reference counterpart: the coarse-space correction of the
	// preconditioner, M^-1 r = V(r) + Phi (Phi^T A Phi)^-1 Phi^T r with Phi the
	// free-cell indicator and up to three fields, and the flexible beta
	const coarseFields = [];
	function setCoarseSpace( freeIndicator = true, fields = [] ) {

		const handles = fields.map( ( f ) => f.handle );
		addon.pressure3SetCoarseSpace( solver, freeIndicator ? 1 : 0, handles );
		coarseFields.length = 0;
		coarseFields.push( ...fields );

	}

	function setFlexible( on = true ) {

		addon.pressure3SetFlexible( solver, on ? 1 : 0 );

	}

	return {
		project, pressure, b: view( FIELD_RHS, 'float' ), diagnostics, settings,
		dirichletMask: view( FIELD_MASK, 'u8' ), ventMask: view( FIELD_VENT, 'u8' ),
		setCoarseSpace, setFlexible,
		// the native solver reads these fields on every project(): hold them
		inputs: keep, coarseFields
	};

}

// ----------------------------------------------------- seeding, transfers
export function createJitterRandom( seed ) {

	// mulberry32 (jitter_random.js:29-42)
	let state = seed >>> 0;
	return function random() {

		state = ( state + 0x6D2B79F5 ) >>> 0;
		let t = state;
		t = Math.imul( t ^ ( t >>> 15 ), t | 1 );
		t ^= t + Math.imul( t ^ ( t >>> 7 ), t | 61 );
		return ( ( t ^ ( t >>> 14 ) ) >>> 0 ) / 4294967296;

	};

}

export function computeFlipBoxSeed3( { boxMin, boxMax, gridSpacingX, gridSpacingY, gridSpacingZ, particlesPerCellAxis = 2, jitter = 0.2, randomSeed = DEFAULT_JITTER_SEED } ) {

	const cells = [ 0, 1, 2 ].map( ( a ) => Math.max( 1, Math.round( ( boxMax[ a ] - boxMin[ a ] ) / [ gridSpacingX, gridSpacingY, gridSpacingZ ][ a ] ) ) );
	const h = [ gridSpacingX, gridSpacingY, gridSpacingZ ];
	const ppa = particlesPerCellAxis;
	const sub = h.map( ( x ) => x / ppa );
	const count = cells[ 0 ] * cells[ 1 ] * cells[ 2 ] * ppa * ppa * ppa;
	const positionsArray = new Float32Array( count * 3 );
	const random = createJitterRandom( randomSeed );
	let q = 0;
	for ( let ck = 0; ck < cells[ 2 ]; ck ++ ) for ( let cj = 0; cj < cells[ 1 ]; cj ++ ) for ( let ci = 0; ci < cells[ 0 ]; ci ++ ) {

		for ( let sk = 0; sk < ppa; sk ++ ) for ( let sj = 0; sj < ppa; sj ++ ) for ( let si = 0; si < ppa; si ++ ) {

			const jx = ( random() * 2 - 1 ) * jitter * sub[ 0 ];
			const jy = ( random() * 2 - 1 ) * jitter * sub[ 1 ];
			const jz = ( random() * 2 - 1 ) * jitter * sub[ 2 ];
			positionsArray[ q ++ ] = boxMin[ 0 ] + ci * h[ 0 ] + ( si + 0.5 ) * sub[ 0 ] + jx;
			positionsArray[ q ++ ] = boxMin[ 1 ] + cj * h[ 1 ] + ( sj + 0.5 ) * sub[ 1 ] + jy;
			positionsArray[ q ++ ] = boxMin[ 2 ] + ck * h[ 2 ] + ( sk + 0.5 ) * sub[ 2 ] + jz;

		}

	}

	return { count, positionsArray, velocitiesArray: new Float32Array( count * 3 ) };

}

function gridArg( grid ) {

	return { n: grid.resolution, h: grid.gridSpacing, origin: grid.origin };

}

// cellIndexOf + stable sort by cell: { cell, perm, cellStart, outside }
export function binParticles( grid, positions, count ) {

	const n = grid.resolution;
	const bins = {
		cell: arrayN( 'int', Math.max( count, 1 ) ), perm: arrayN( 'int', Math.max( count, 1 ) ),
		cellStart: arrayN( 'int', n[ 0 ] * n[ 1 ] * n[ 2 ] + 1 )
	};
	bins.outside = addon.binParticles( gridArg( grid ), positions.handle, count, bins.cell.handle, bins.perm.handle, bins.cellStart.handle );
	return bins;

}

export function particleToGrid( grid, positions, velocities, bins, weightEpsilon = 1e-4 ) {

	const valid = [ grid.dataSizeU, grid.dataSizeV, grid.dataSizeW ].map( ( s ) => arrayN( 'int', s ) );
	const data = [ grid.dataU, grid.dataV, grid.dataW ];
	for ( let c = 0; c < 3; c ++ ) addon.p2gGather( gridArg( grid ), positions.handle, velocities.handle, bins.perm.handle, bins.cellStart.handle, c, weightEpsilon, data[ c ].handle, valid[ c ].handle );
	return valid;

}

export function gridToParticle( grid, oldGrid, positions, velocities, count, flipRatio = 0.97, velocityDamping = 0.02 ) {

	addon.g2p( gridArg( grid ), positions.handle, velocities.handle, count,
		[ grid.dataU.handle, grid.dataV.handle, grid.dataW.handle ],
		[ oldGrid.dataU.handle, oldGrid.dataV.handle, oldGrid.dataW.handle ], flipRatio, velocityDamping );

}

export function advectParticles( grid, positions, count, dt, maxSubsteps = 32 ) {

	addon.advectParticles( gridArg( grid ), positions.handle, count, [ grid.dataU.handle, grid.dataV.handle, grid.dataW.handle ], dt, maxSubsteps );

}

export function advectFaceCentered3( input, output, dt, maxSubsteps = 32 ) {

	addon.advectFaces( gridArg( input ), [ input.dataU.handle, input.dataV.handle, input.dataW.handle ],
		[ output.dataU.handle, output.dataV.handle, output.dataW.handle ], dt, maxSubsteps );

}

// createCellCenteredScalarGrid3 (grid_data3.js:181-218): cell data at
// origin + h/2
export function createCellCenteredScalarGrid3( nx, ny, nz, hx, hy, hz, ox = 0, oy = 0, oz = 0 ) {

	const grid = {
		resolution: [ nx, ny, nz ], gridSpacing: [ hx, hy, hz ], origin: [ ox, oy, oz ],
		dataSize: [ nx, ny, nz ], dataOrigin: [ ox + hx * 0.5, oy + hy * 0.5, oz + hz * 0.5 ],
		data: arrayN( 'float', [ nx, ny, nz ] ), cellCentered: true
	};
	grid.clear = () => grid.data.fill( 0 );
	return grid;

}

// createSemiLagrangianAdvectionSolver3 (advection_solver3.js:67-362), no
// collider: velocityGrid is bound here (its contents are read at each
// dispatch); dt is a number or an array0('float') field read at dispatch;
// order 1 or 2 (MacCormack).  advectFaceCentered3 / advectScalar3 return
// 0-arg dispatchers, as the reference's do.
export function createSemiLagrangianAdvectionSolver3( { velocityGrid, collider, dt, maxSubsteps = 32, order = 1 } = {} ) {

	if ( ! velocityGrid ) throw new Error( 'createSemiLagrangianAdvectionSolver3: options.velocityGrid is required.' );
	if ( collider ) {

		const cr = collider.grid.resolution, vr = velocityGrid.resolution;
		if ( cr[ 0 ] !== vr[ 0 ] || cr[ 1 ] !== vr[ 1 ] || cr[ 2 ] !== vr[ 2 ] ) {

			throw new Error( 'createSemiLagrangianAdvectionSolver3: the collider\'s grid must be the velocity grid\'s cells.' );

		}

	}

	if ( order !== 1 && order !== 2 ) throw new Error( 'createSemiLagrangianAdvectionSolver3: order must be 1 or 2.' );
	const readDt = () => {

		if ( typeof dt === 'number' ) return dt;
		if ( dt && dt.handle ) {

			const out = new Float32Array( 1 );
			addon.fieldDownload( dt.handle, out );
			return out[ 0 ];

		}

		throw new Error( 'createSemiLagrangianAdvectionSolver3: options.dt must be a number or an array0(\'float\') field.' );

	};

	const handle = addon.advection3Create( gridArg( velocityGrid ),
		[ velocityGrid.dataU.handle, velocityGrid.dataV.handle, velocityGrid.dataW.handle ], readDt(), maxSubsteps, order );
	if ( collider ) addon.advection3SetCollider( handle, collider.grid.data.handle );
	const live = typeof dt !== 'number';
	function advectFaceCentered3( input, output ) {

		return () => {

			if ( live ) addon.advection3SetDt( handle, readDt() );
			addon.advection3Faces( handle, [ input.dataU.handle, input.dataV.handle, input.dataW.handle ],
				[ output.dataU.handle, output.dataV.handle, output.dataW.handle ] );

		};

	}

	function advectScalar3( input, output ) {

		for ( const g of [ input, output ] ) {

			if ( ! g.cellCentered ) throw new Error( 'advectScalar3: input and output must be cell-centred scalar grids (createCellCenteredScalarGrid3).' );

		}

		return () => {

			if ( live ) addon.advection3SetDt( handle, readDt() );
			addon.advection3Scalar( handle, input.data.handle, output.data.handle );

		};

	}

	return { advectFaceCentered3, advectScalar3 };

}

export function twoPhaseBeta( grid, bins, gasDensity, liquidDensity, fraction ) {

	const beta = { u: arrayN( 'float', grid.dataSizeU ), v: arrayN( 'float', grid.dataSizeV ), w: arrayN( 'float', grid.dataSizeW ) };
	addon.twoPhaseBeta( gridArg( grid ), bins ? bins.cellStart.handle : null, gasDensity, liquidDensity, fraction.handle, [ beta.u.handle, beta.v.handle, beta.w.handle ] );
	return beta;

}

// ------------------------------------------------------ FLIP grid stages
// createExtrapolateToRegion3 (array_utils3.js:166-228): run(numberOfIterations = 5)
export function createExtrapolateToRegion3( inputField, validField, outputField, shape = validField.shape ) {

	if ( inputField.count !== validField.count || outputField.count !== validField.count ) {

		throw new Error( `createExtrapolateToRegion3: shape mismatch [${ shape }].` );

	}

	const handle = addon.extrapolatorCreate( inputField.handle, validField.handle, outputField.handle );
	function run( numberOfIterations = 5 ) {

		addon.extrapolatorRun( handle, numberOfIterations );

	}

	run.inputs = { inputField, validField, outputField };
	return run;

}

// primitive_sdf3.js:21-70 shape descriptors
export function sphere( center, radius ) {

	return { type: 'sphere', center, radius };

}

export function box( center, halfExtents, rotation = null ) {

	return { type: 'box', center, halfExtents, rotation };

}

// createSDFStaticCollider3 (sdf_collider3.js:24-262): the SDF lives on a
// cell-centred grid; computeFaceWeights evaluates the shapes analytically
// (FP64) on a velocity grid's faces.
export function createSDFStaticCollider3( resolutionX, resolutionY, resolutionZ, gridSpacingX, gridSpacingY, gridSpacingZ, originX = 0, originY = 0, originZ = 0 ) {

	const resolution = [ resolutionX, resolutionY, resolutionZ ];
	const gridSpacing = [ gridSpacingX, gridSpacingY, gridSpacingZ ];
	const origin = [ originX, originY, originZ ];
	const data = arrayN( 'float', resolution );
	let currentShapes = [];
	let currentInvert = false;
	const collider = { grid: { resolution, gridSpacing, origin, data }, frictionCoefficient: 0.5 };
	collider.addShapes = ( shapes, { invert = false } = {} ) => {

		currentShapes = shapes;
		currentInvert = invert;
		addon.colliderSdf( { n: resolution, h: gridSpacing, origin }, shapes, invert, data.handle );

	};

	collider.addShape = ( shape, options ) => collider.addShapes( [ shape ], options );
	collider.computeFaceWeights = ( nx, ny, nz, hx, hy, hz, ox = 0, oy = 0, oz = 0 ) => {

		const u = arrayN( 'float', [ nx + 1, ny, nz ] );
		const v = arrayN( 'float', [ nx, ny + 1, nz ] );
		const w = arrayN( 'float', [ nx, ny, nz + 1 ] );
		addon.colliderFaceWeights( { n: [ nx, ny, nz ], h: [ hx, hy, hz ], origin: [ ox, oy, oz ] }, currentShapes, currentInvert,
			[ u.handle, v.handle, w.handle ] );
		return { u, v, w, fields: { u, v, w } };

	};

	return collider;

}

// createSDFInflow3 / createSDFOutflow3 (sdf_inflow_outflow3.js:15-30): SDF
// objects on their own grid (the velocity grid's cells), an inflow with its
// velocity [vx, vy, vz] and mode 'set' or 'add'
export function createSDFInflow3( resolutionX, resolutionY, resolutionZ, gridSpacingX, gridSpacingY, gridSpacingZ,
	originX = 0, originY = 0, originZ = 0, { velocity = [ 0, 0, 0 ], mode = 'set' } = {} ) {

	if ( mode !== 'set' && mode !== 'add' ) throw new Error( `createSDFInflow3: mode must be 'set' or 'add', got '${ mode }'.` );
	const sdf = createSDFStaticCollider3( resolutionX, resolutionY, resolutionZ, gridSpacingX, gridSpacingY, gridSpacingZ, originX, originY, originZ );
	return { ...sdf, velocity, mode };

}

export function createSDFOutflow3( resolutionX, resolutionY, resolutionZ, gridSpacingX, gridSpacingY, gridSpacingZ,
	originX = 0, originY = 0, originZ = 0 ) {

	return createSDFStaticCollider3( resolutionX, resolutionY, resolutionZ, gridSpacingX, gridSpacingY, gridSpacingZ, originX, originY, originZ );

}

// the SDF's collocatedValueAtPosition3 on the host (trilinear, clamped
// indices), from a copy read once
function hostSampler( sdfObject ) {

	let data = null;
	const { resolution: n, gridSpacing: h, origin: o } = sdfObject.grid;
	return ( pos ) => {

		if ( ! data ) data = readFloats( sdfObject.grid.data );
		const idx = [], f = [];
		for ( let a = 0; a < 3; a ++ ) {

			const g = ( pos[ a ] - ( o[ a ] + 0.5 * h[ a ] ) ) / h[ a ];
			const i0 = Math.floor( g );
			f.push( g - i0 );
			idx.push( [ Math.min( Math.max( i0, 0 ), n[ a ] - 1 ), Math.min( Math.max( i0 + 1, 0 ), n[ a ] - 1 ) ] );

		}

		let acc = 0;
		for ( let c = 0; c < 8; c ++ ) {

			const bx = c & 1, by = ( c >> 1 ) & 1, bz = ( c >> 2 ) & 1;
			const wgt = ( bx ? f[ 0 ] : 1 - f[ 0 ] ) * ( by ? f[ 1 ] : 1 - f[ 1 ] ) * ( bz ? f[ 2 ] : 1 - f[ 2 ] );
			acc += data[ idx[ 0 ][ bx ] + n[ 0 ] * ( idx[ 1 ][ by ] + n[ 1 ] * idx[ 2 ][ bz ] ) ] * wgt;

		}

		return acc;

	};

}

function readFloats( field ) {

	const out = new Float32Array( field.count );
	addon.fieldDownload( field.handle, out );
	return out;

}

// createOutflowPressureDirichlet3( outflows, { target } ) and
// combineDirichlet3( a, b ) (sdf_inflow_outflow3.js:44-94): dirichlet(pos)
// functions of the reference's shape; with target 0 the pressure solver
// evaluates an outflow condition (combined with any other as `a`) on the
// GPU at its cell centres instead of calling the function per cell.
export function createOutflowPressureDirichlet3( outflows, { target = 0 } = {} ) {

	const list = Array.isArray( outflows ) ? outflows : [ outflows ];
	const samplers = list.map( hostSampler );
	const dirichlet = ( pos ) => ( { active: samplers.some( ( sample ) => sample( pos ) < 0 ), target } );
	if ( target === 0 ) {

		dirichlet.outflows = list;
		dirichlet.evaluateFields = ( resolution, gridSpacing, origin, user = null ) => {

			const active = arrayN( 'u8', resolution ), tgt = arrayN( 'float', resolution );
			addon.outflow3Dirichlet( { n: resolution, h: gridSpacing, origin }, list.map( ( o ) => o.grid.data.handle ),
				user ? user.active.handle : null, user && user.target ? user.target.handle : null, active.handle, tgt.handle );
			return { active, target: tgt };

		};

	}

	return dirichlet;

}

export function combineDirichlet3( a, b ) {

	if ( ! a ) return b;
	if ( ! b ) return a;
	const dirichlet = ( pos ) => {

		const ra = a( pos ), rb = b( pos );
		return { active: ra.active || rb.active, target: rb.active ? rb.target : ra.target };

	};

	if ( a.outflows ) {

		dirichlet.evaluateFields = ( resolution, gridSpacing, origin ) => {

			let user;
			if ( b.evaluateFields ) user = b.evaluateFields( resolution, gridSpacing, origin );
			else {

				const { active, target } = evaluateDirichlet( b, resolution, gridSpacing, origin );
				user = { active: arrayN( 'u8', resolution ).fromArray( active ), target: arrayN( 'float', resolution ).fromArray( target ) };

			}

			return a.evaluateFields( resolution, gridSpacing, origin, user );

		};

	}

	return dirichlet;

}

// createGridBlockedBoundaryConditionSolver3 (grid_blocked_boundary_condition_solver3.js:47-490)
// for a static collider and createSDFInflow3 objects (their SDF on the
// velocity grid's cells, at most 4)
export function createGridBlockedBoundaryConditionSolver3( velocity, resolutionX, resolutionY, resolutionZ,
	gridSpacingX, gridSpacingY, gridSpacingZ, originX = 0, originY = 0, originZ = 0,
	colliderSDF = null, inflows = null, colliderFaceWeights = null ) {

	const grid = { n: [ resolutionX, resolutionY, resolutionZ ], h: [ gridSpacingX, gridSpacingY, gridSpacingZ ], origin: [ originX, originY, originZ ] };
	const weights = colliderFaceWeights ? [ colliderFaceWeights.fields.u.handle, colliderFaceWeights.fields.v.handle, colliderFaceWeights.fields.w.handle ] : null;
	const handle = addon.boundary3Create( grid, [ velocity.dataU.handle, velocity.dataV.handle, velocity.dataW.handle ],
		colliderSDF ? colliderSDF.grid.data.handle : null, weights, colliderSDF ? colliderSDF.frictionCoefficient : 0.5 );
	const inflowList = ! inflows ? [] : ( Array.isArray( inflows ) ? inflows : [ inflows ] );
	for ( const inflow of inflowList ) {

		addon.boundary3AddInflow( handle, inflow.grid.data.handle, inflow.velocity, inflow.mode === 'add' ? 1 : 0 );

	}

	let flag = 63;
	const solver = {
		velocity, collider: colliderSDF, inflows, inputs: { colliderFaceWeights, inflowList },
		get closedDomainBoundaryFlag() {

			return flag;

		},
		set closedDomainBoundaryFlag( f ) {

			addon.boundary3SetClosed( handle, f );
			flag = f;

		},
		constrainVelocity( extrapolationDepth = 5 ) {

			addon.boundary3Constrain( handle, extrapolationDepth );

		}
	};
	return solver;

}

// createGridFlipSolver3 (grid_flip_solver3.js:135-1155).  The solver owns its
// particle and velocity fields: `velocityGrid` supplies the resolution,
// spacing and origin, and solver.velocityGrid exposes the solver's own
// dataU/V/W.  `collider` is a createSDFStaticCollider3 object.
export function createGridFlipSolver3( {
	velocityGrid, gridSpacing, origin, maxParticles, dt, gravity = [ 0, - 9.81, 0 ],
	reducedPressure = true, flipRatio = 0.97, velocityDamping = 0.02, p2gAtomicScale = 65536,
	weightEpsilon = 1e-4, closedDomainBoundaryFlag, collider, colliderFaceWeights, resample = {}, pressure = {},
	carryConcentration = false, mixing = 0, fade = 0, concentrationAtomicScale = 4096, ambientDensity,
	componentDensity, massWeightedTransfer = false, pressurePin = null, colliderPushThresh = 0, colliderPushShift = 0
} = {} ) {

	if ( ! velocityGrid ) throw new Error( 'createGridFlipSolver3: options.velocityGrid is required.' );
	if ( ! maxParticles ) throw new Error( 'createGridFlipSolver3: options.maxParticles is required.' );
	const { enabled: resampleEnabled = true, minParticlesPerCell = 4, maxParticlesPerCell = 8, phaseAware = false } = resample;
	let pressurePinMode = 0;
	let pinCell = null;
	if ( pressurePin === 'topCenter' ) pressurePinMode = 1;
	else if ( Array.isArray( pressurePin ) && pressurePin.length === 3 ) {

		pressurePinMode = 2;
		pinCell = pressurePin;

	} else if ( pressurePin !== null && pressurePin !== undefined ) {

		throw new Error( `createGridFlipSolver3: pressurePin must be 'topCenter', null, or an [i, j, k] cell index, got ${ JSON.stringify( pressurePin ) }.` );

	}

	const coupled = ambientDensity !== undefined && componentDensity !== undefined;
	const resolution = velocityGrid.resolution.slice();
	const spacing = gridSpacing || velocityGrid.gridSpacing;
	const org = origin || velocityGrid.origin;
	const pre = pressure.preconditioner || 'multigrid';
	if ( ! Object.prototype.hasOwnProperty.call( PRECONDITIONERS, pre ) ) {

		throw new Error( `grid_pressure_solver3: unknown preconditioner '${ pre }' (expected ${ Object.keys( PRECONDITIONERS ).join( ', ' ) }).` );

	}

	const handle = addon.flip3Create( {
		resolution, gridSpacing: spacing, origin: org, maxParticles, dt, gravity, reducedPressure, flipRatio,
		velocityDamping, p2gAtomicScale, weightEpsilon, closedDomainBoundaryFlag, resampleEnabled,
		minParticlesPerCell, maxParticlesPerCell,
		colliderSdf: collider ? collider.grid.data.handle : null,
		colliderWeights: colliderFaceWeights ? [ colliderFaceWeights.fields.u.handle, colliderFaceWeights.fields.v.handle, colliderFaceWeights.fields.w.handle ] : null,
		friction: collider ? collider.frictionCoefficient : 0.5,
		multigrid: pressure.multigrid || {}, preconditionerKind: PRECONDITIONERS[ pre ],
		tolerance: pressure.tolerance === undefined ? 1e-6 : pressure.tolerance,
		maxIterations: pressure.maxIterations === undefined ? null : pressure.maxIterations,
		maxPlausiblePressure: pressure.maxPlausiblePressure === undefined ? null : pressure.maxPlausiblePressure,
		carryConcentration, mixing, fade, concentrationAtomicScale,
		ambientDensity: coupled ? ambientDensity : null, componentDensity: coupled ? componentDensity : null,
		massWeightedTransfer, pressurePinMode, pinCell, phaseAware, colliderPushThresh, colliderPushShift
	} );
	const view = ( which, type, shape ) => wrapField( addon.flip3Field( handle, which ), type, shape );
	const [ nx, ny, nz ] = resolution;
	const solverGrid = {
		resolution, gridSpacing: spacing, origin: org,
		dataSizeU: [ nx + 1, ny, nz ], dataSizeV: [ nx, ny + 1, nz ], dataSizeW: [ nx, ny, nz + 1 ],
		dataU: view( 3, 'float', [ nx + 1, ny, nz ] ), dataV: view( 4, 'float', [ nx, ny + 1, nz ] ),
		dataW: view( 5, 'float', [ nx, ny, nz + 1 ] )
	};
	const diagnostics = { converged: null, rejected: false, iterations: null, stoppedBy: null };
	async function onAdvanceTimeStep() {

		const d = await addon.flip3StepAsync( handle );
		Object.assign( diagnostics, d, { stoppedBy: d.stoppedBy === 'converged' ? 'none' : d.stoppedBy } );
		return diagnostics;

	}

	// the step's dt from the next step on (the reference's dt is a live node
	// that createGridAdaptiveTimeStep3 writes)
	function setDt( value ) {

		addon.flip3SetDt( handle, value );

	}

	// the free-cell coarse correction of the step's projection (no reference
	// counterpart)
	function setCoarseSpace( freeIndicator = true ) {

		addon.flip3SetCoarseSpace( handle, freeIndicator ? 1 : 0 );

	}

	return {
		onAdvanceTimeStep, setDt, setCoarseSpace, settings: { batchStages: true },
		positions: view( 0, 'vec3', [ maxParticles ] ), velocities: view( 1, 'vec3', [ maxParticles ] ),
		fluidMask: view( 2, 'float', resolution ), pressure: { data: view( 6, 'float', resolution ) },
		velocityGrid: solverGrid, diagnostics, inputs: { collider, colliderFaceWeights },
		concentration: carryConcentration ? view( 7, 'float', [ maxParticles ] ) : null,
		cellConcentration: carryConcentration ? view( 8, 'float', resolution ) : null,
		cellDensity: coupled ? view( 9, 'float', resolution ) : null
	};

}

// createGridSmokeSolver3 (grid_smoke_solver3.js:25-154), closed walls, no
// collider/inflow/outflow objects.  The solver owns its velocity and scalar
// fields: velocityGrid.dataU/V/W, density.current, temperature.current.
// sourceSphere(center, radius, density, temperature) re-injects a ball with
// max, as examples/34-smoke-plume-3d/main.js:146-158 does.
export function createGridSmokeSolver3( {
	velocityGrid, gridSpacing, origin, dt, closedDomainBoundaryFlag,
	buoyancySmokeDensityFactor = - 0.000625, buoyancyTemperatureFactor = 5.0, ambientTemperature = 0,
	smokeDecay = 0.001, temperatureDecay = 0.001, up = [ 0, 1, 0 ], collider, inflows, outflows, pressure = {},
	advection = {}
} = {} ) {

	if ( ! velocityGrid ) throw new Error( 'createGridSmokeSolver3: options.velocityGrid is required.' );
	if ( collider || inflows || outflows ) throw new Error( 'createGridSmokeSolver3: collider, inflow and outflow objects are not supported by this build.' );
	const resolution = velocityGrid.resolution.slice();
	const spacing = gridSpacing || velocityGrid.gridSpacing;
	const org = origin || velocityGrid.origin;
	const pre = pressure.preconditioner || 'multigrid';
	if ( ! Object.prototype.hasOwnProperty.call( PRECONDITIONERS, pre ) ) {

		throw new Error( `grid_pressure_solver3: unknown preconditioner '${ pre }' (expected ${ Object.keys( PRECONDITIONERS ).join( ', ' ) }).` );

	}

	const handle = addon.smoke3Create( {
		resolution, gridSpacing: spacing, origin: org, dt, closedDomainBoundaryFlag, up,
		buoyancySmokeDensityFactor, buoyancyTemperatureFactor, ambientTemperature, smokeDecay, temperatureDecay,
		multigrid: pressure.multigrid || {}, preconditionerKind: PRECONDITIONERS[ pre ],
		tolerance: pressure.tolerance === undefined ? 1e-6 : pressure.tolerance,
		maxIterations: pressure.maxIterations === undefined ? null : pressure.maxIterations,
		advectionOrder: advection.order === undefined ? 1 : advection.order,
		maxSubsteps: advection.maxSubsteps === undefined ? 32 : advection.maxSubsteps
	} );
	const [ nx, ny, nz ] = resolution;
	const view = ( which, shape ) => wrapField( addon.smoke3Field( handle, which ), 'float', shape );
	const grid = {
		resolution, gridSpacing: spacing, origin: org,
		dataSizeU: [ nx + 1, ny, nz ], dataSizeV: [ nx, ny + 1, nz ], dataSizeW: [ nx, ny, nz + 1 ],
		dataU: view( 0, [ nx + 1, ny, nz ] ), dataV: view( 1, [ nx, ny + 1, nz ] ), dataW: view( 2, [ nx, ny, nz + 1 ] )
	};
	const diagnostics = { converged: null, rejected: false, iterations: null, stoppedBy: null };
	async function onAdvanceTimeStep() {

		const d = await addon.smoke3StepAsync( handle );
		Object.assign( diagnostics, d, { stoppedBy: d.stoppedBy === 'converged' ? 'none' : d.stoppedBy } );
		return diagnostics;

	}

	return {
		onAdvanceTimeStep, velocityGrid: grid, diagnostics,
		density: { get current() { return view( 3, resolution ); } },
		temperature: { get current() { return view( 4, resolution ); } },
		sourceSphere: ( center, radius, density = 1, temperature = 1 ) => addon.smoke3Source( handle, center, radius, density, temperature )
	};

}

// ------------------------------------------------------------------- 2-D
export const DIRECTION_NONE = 0, DIRECTION_LEFT = 1, DIRECTION_RIGHT = 2, DIRECTION_DOWN = 4, DIRECTION_UP = 8;
export const DIRECTION_ALL = 15;

// createFaceCenteredGrid2 (grid_data2.js:161-198)
export function createFaceCenteredGrid2( nx, ny, hx, hy, ox = 0, oy = 0 ) {

	const dataSizeU = [ nx + 1, ny ];
	const dataSizeV = [ nx, ny + 1 ];
	const grid = {
		resolution: [ nx, ny ], gridSpacing: [ hx, hy ], origin: [ ox, oy ], dataSizeU, dataSizeV,
		dataOriginU: [ ox, oy + hy * 0.5 ], dataOriginV: [ ox + hx * 0.5, oy ],
		dataU: arrayN( 'float', dataSizeU ), dataV: arrayN( 'float', dataSizeV )
	};
	grid.dataU.fill( 0 );
	grid.dataV.fill( 0 );
	grid.clear = () => {

		grid.dataU.fill( 0 );
		grid.dataV.fill( 0 );

	};

	return grid;

}

// createSDFStaticCollider2 (sdf_collider2.js:24-130): the SDF lives in a
// cell-centred field (0 until polygons are added, as the reference's zeroed
// grid); addPolygons rasterises it on the GPU in FP64.
export function createSDFStaticCollider2( resolutionX, resolutionY, gridSpacingX, gridSpacingY, originX = 0, originY = 0 ) {

	const resolution = [ resolutionX, resolutionY ];
	const gridSpacing = [ gridSpacingX, gridSpacingY ];
	const origin = [ originX, originY ];
	const data = arrayN( 'float', resolution ).fill( 0 );
	const collider = { grid: { resolution, gridSpacing, origin, data }, frictionCoefficient: 0.5 };
	collider.addPolygons = ( polygons, { invert = false } = {} ) => {

		addon.polygonSdf2( data.handle, { resolution, gridSpacing, origin }, polygons, invert );

	};

	collider.addPolygon = ( points, options ) => collider.addPolygons( [ points ], options );
	return collider;

}

// createSDFInflow2 / createSDFOutflow2 (sdf_inflow_outflow2.js:102-120)
export function createSDFInflow2( resolutionX, resolutionY, gridSpacingX, gridSpacingY, originX = 0, originY = 0, options = {} ) {

	const { velocity = [ 0, 0 ], mode = 'set' } = options;
	if ( mode !== 'set' && mode !== 'add' ) throw new Error( `createSDFInflow2: mode must be 'set' or 'add'.` );
	return { ...createSDFStaticCollider2( resolutionX, resolutionY, gridSpacingX, gridSpacingY, originX, originY ), velocity, mode };

}

export function createSDFOutflow2( resolutionX, resolutionY, gridSpacingX, gridSpacingY, originX = 0, originY = 0 ) {

	return createSDFStaticCollider2( resolutionX, resolutionY, gridSpacingX, gridSpacingY, originX, originY );

}

// computeAdaptiveSubSteps (time/cfl.js:38-48)
export function computeAdaptiveSubSteps( maxVelocityMagnitude, minGridSpacing, frameDt, options = {} ) {

	const c = options.courantNumber === undefined ? 5 : options.courantNumber;
	const m = options.maxSubSteps === undefined ? 32 : options.maxSubSteps;
	return addon.adaptiveSubSteps( maxVelocityMagnitude, minGridSpacing, frameDt, c, m );

}

// createMaxAbsReducer (reduction.js:43-77)
export function createMaxAbsReducer( fields ) {

	const list = Array.isArray( fields ) ? fields : [ fields ];
	return { read: async () => addon.maxAbs( list.map( ( f ) => f.handle ) ) };

}

function asList( x ) {

	return ! x ? [] : ( Array.isArray( x ) ? x : [ x ] );

}

// createGridSolver2 (grid_solver2.js:131-255) without external forces.  The
// solver owns its velocity fields (velocityGrid.dataU/V are the solver's;
// the grid passed supplies resolution, spacing and origin, and its current
// contents are copied in).  dt may be a number or an array0('float') field
// (a live dt, re-read every step, as createGridAdaptiveTimeStep2 writes it).
export function createGridSolver2( {
	velocityGrid, gridSpacing, origin, force, dirichlet, collider, inflows, outflows, outflowVelocityBC = true,
	closedDomainBoundaryFlag, dt, advection = {}, pressure = {}
} = {} ) {

	if ( ! velocityGrid ) throw new Error( 'createGridSolver2: options.velocityGrid is required.' );
	if ( force ) throw new Error( 'createGridSolver2: external forces are not supported by this build.' );
	if ( advection.order !== undefined && advection.order !== 1 && advection.order !== 2 ) throw new Error( 'createGridSolver2: advection.order must be 1 or 2.' );
	if ( advection.collider && advection.collider !== collider ) throw new Error( 'createGridSolver2: advection.collider must be the solver\'s collider.' );
	const resolution = velocityGrid.resolution.slice();
	const spacing = gridSpacing || velocityGrid.gridSpacing;
	const org = origin || velocityGrid.origin;
	const pre = pressure.preconditioner || 'multigrid';
	if ( ! Object.prototype.hasOwnProperty.call( PRECONDITIONERS, pre ) ) {

		throw new Error( `grid_pressure_solver2: unknown preconditioner '${ pre }' (expected ${ Object.keys( PRECONDITIONERS ).join( ', ' ) }).` );

	}

	let dirichletActive = null, dirichletTarget = null;
	if ( dirichlet ) {

		if ( typeof dirichlet === 'function' ) {

			// dirichlet(pos) at the cell centres (grid_pressure_solver2.js:403-417)
			const [ nx, ny ] = resolution;
			const act = new Uint8Array( nx * ny );
			const tgt = new Float32Array( nx * ny );
			for ( let j = 0; j < ny; j ++ ) for ( let i = 0; i < nx; i ++ ) {

				const r = dirichlet( [ org[ 0 ] + ( i + 0.5 ) * spacing[ 0 ], org[ 1 ] + ( j + 0.5 ) * spacing[ 1 ] ] );
				act[ i + nx * j ] = r.active ? 1 : 0;
				tgt[ i + nx * j ] = r.target || 0;

			}

			dirichletActive = arrayN( 'u8', resolution ).fromArray( act );
			dirichletTarget = arrayN( 'float', resolution ).fromArray( tgt );

		} else {

			dirichletActive = dirichlet.active;
			dirichletTarget = dirichlet.target;

		}

	}

	const handle = addon.solver2Create( {
		resolution, gridSpacing: spacing, origin: org,
		colliderSdf: collider ? collider.grid.data.handle : null,
		friction: collider ? collider.frictionCoefficient : 0.5,
		advectCollider: !! advection.collider, maxSubsteps: advection.maxSubsteps === undefined ? 32 : advection.maxSubsteps,
		advectionOrder: advection.order === undefined ? 1 : advection.order,
		inflows: asList( inflows ).map( ( q ) => ( { sdf: q.grid.data.handle, velocity: q.velocity, mode: q.mode === 'add' ? 1 : 0 } ) ),
		outflows: asList( outflows ).map( ( q ) => q.grid.data.handle ), outflowVelocityBC,
		dirichletActive: dirichletActive ? dirichletActive.handle : null,
		dirichletTarget: dirichletTarget ? dirichletTarget.handle : null,
		closedDomainBoundaryFlag,
		multigrid: pressure.multigrid || {}, preconditionerKind: PRECONDITIONERS[ pre ],
		tolerance: pressure.tolerance === undefined ? 1e-6 : pressure.tolerance,
		maxIterations: pressure.maxIterations === undefined ? null : pressure.maxIterations
	} );
	const keep = [ collider, inflows, outflows, dirichletActive, dirichletTarget ];
	const [ nx, ny ] = resolution;
	const view = ( which, shape ) => wrapField( addon.solver2Field( handle, which ), 'float', shape );
	const grid = {
		resolution, gridSpacing: spacing, origin: org, dataSizeU: [ nx + 1, ny ], dataSizeV: [ nx, ny + 1 ],
		dataU: view( 0, [ nx + 1, ny ] ), dataV: view( 1, [ nx, ny + 1 ] )
	};
	if ( velocityGrid.dataU && velocityGrid.dataU.handle ) {

		addon.fieldCopy( grid.dataU.handle, velocityGrid.dataU.handle ); // (destination, source)
		addon.fieldCopy( grid.dataV.handle, velocityGrid.dataV.handle );

	}

	const diagnostics = { converged: null, rejected: false, iterations: null, stoppedBy: null };
	const readDt = () => {

		if ( typeof dt === 'number' ) return dt;
		if ( dt && dt.handle ) {

			const out = new Float32Array( 1 );
			addon.fieldDownload( dt.handle, out );
			return out[ 0 ];

		}

		throw new Error( 'createGridSolver2: options.dt must be a number or an array0(\'float\') field.' );

	};

	async function onAdvanceTimeStep() {

		const d = await addon.solver2StepAsync( handle, readDt() );
		Object.assign( diagnostics, d, { stoppedBy: d.stoppedBy === 'converged' ? 'none' : d.stoppedBy } );
		return diagnostics;

	}

	return {
		onAdvanceTimeStep, velocityGrid: grid, diagnostics, keep,
		get pressure() { return { data: view( 4, resolution ) }; }
	};

}

// createGridAdaptiveTimeStep2 (grid_adaptive_timestep2.js:15-44): update()
// reads max |u| over the face arrays, writes the sub-step dt into the live
// dt field and returns the sub-step count.
export function createGridAdaptiveTimeStep2( { velocityGrid, gridSpacing, dt, targetDt, courantNumber, maxSubSteps } ) {

	if ( typeof dt === 'number' || ! dt || ! dt.handle ) {

		throw new Error( 'createGridAdaptiveTimeStep2: options.dt must be a live array0(\'float\') field.' );

	}

	const reducer = createMaxAbsReducer( [ velocityGrid.dataU, velocityGrid.dataV ] );
	const minGridSpacing = Math.min( ...gridSpacing );
	const state = { lastNumSubSteps: 1, lastSubDt: targetDt };
	async function update() {

		const maxVel = await reducer.read();
		const { numSubSteps, subDt } = computeAdaptiveSubSteps( maxVel, minGridSpacing, targetDt, { courantNumber, maxSubSteps } );
		dt.fromArray( new Float32Array( [ subDt ] ) );
		state.lastNumSubSteps = numSubSteps;
		state.lastSubDt = subDt;
		return numSubSteps;

	}

	return { update, state };

}

// ------------------------------------------------------------- fixtures
// The pressure-system fixture format of export_system.mjs:65-126: meta.json
// and one little-endian float32 .bin per array (b, pressure, dirichletMask,
// ventMask, weightU/V/W), first index fastest.
export async function exportSystem( directory, solver, { source = 'tsl_array_n_amd', frame = 0, faceWeights = null } = {} ) {

	const fs = require( 'fs' );
	const path = require( 'path' );
	fs.mkdirSync( directory, { recursive: true } );
	const res = solver.pressure.resolution.slice();
	const asFloat = async ( field ) => Float32Array.from( await field.toArray() );
	const arrays = {
		b: await asFloat( solver.b ), pressure: await asFloat( solver.pressure.data ),
		dirichletMask: solver.dirichletMask ? await asFloat( solver.dirichletMask ) : null,
		ventMask: solver.ventMask ? await asFloat( solver.ventMask ) : null,
		weightU: faceWeights ? await asFloat( faceWeights.u ) : null,
		weightV: faceWeights ? await asFloat( faceWeights.v ) : null,
		weightW: faceWeights ? await asFloat( faceWeights.w ) : null
	};
	const written = [];
	for ( const [ name, values ] of Object.entries( arrays ) ) {

		if ( values === null ) continue;
		fs.writeFileSync( path.join( directory, `${ name }.bin` ), Buffer.from( values.buffer ) );
		written.push( name );

	}

	const [ nx, ny, nz ] = res;
	const meta = {
		source, frame, resolution: res, gridSpacing: solver.pressure.gridSpacing,
		dataSizeU: [ nx + 1, ny, nz ], dataSizeV: [ nx, ny + 1, nz ], dataSizeW: [ nx, ny, nz + 1 ],
		diagnostics: { ...solver.diagnostics }, arrays: written
	};
	fs.writeFileSync( path.join( directory, 'meta.json' ), JSON.stringify( meta, null, '\t' ) + '\n' );
	return meta;

}

export function loadSystem( directory ) {

	const fs = require( 'fs' );
	const path = require( 'path' );
	const meta = JSON.parse( fs.readFileSync( path.join( directory, 'meta.json' ), 'utf8' ) );
	const read = ( name ) => {

		const file = path.join( directory, `${ name }.bin` );
		if ( ! fs.existsSync( file ) ) return null;
		const buf = fs.readFileSync( file );
		return new Float32Array( buf.buffer.slice( buf.byteOffset, buf.byteOffset + buf.byteLength ) );

	};

	const arrays = {};
	for ( const name of [ 'b', 'pressure', 'dirichletMask', 'ventMask', 'weightU', 'weightV', 'weightW' ] ) arrays[ name ] = read( name );
	return { meta, resolution: meta.resolution, gridSpacing: meta.gridSpacing, arrays };

}
