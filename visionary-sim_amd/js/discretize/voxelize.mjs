// The voxelisers on the HIP device: voxelizeSdf and voxelizeMesh
// (packages/visionary-sim/src/sim/discretize/lattice/voxelize.js:39-203), with
// the grid layout and helpers the reference exports beside them. The grids
// are the reference's cell for cell (the device evaluates the JS's f64
// expressions in the same order); the layout helpers are host bookkeeping.
import { addon } from '../native.mjs';

/** Empty cells added on every side (voxelize.js:18). */
export const DEFAULT_PADDING = 2;

/** Shape kinds the device evaluates (VS_SHAPE_* in include/vsim.h). */
export const SHAPE_KINDS = Object.freeze({ sphere: 0, box: 1, torus: 2 });

function elementSizeError(elementSize) {
    return new Error(`gridForBounds: elementSize must be a positive length in metres, got ${elementSize}`);
}

/**
 * gridForBounds (voxelize.js:39-67): an empty grid covering `bounds`.
 * @param {{ min: number[], max: number[] }} bounds - Bounds to cover.
 * @param {{ elementSize: number, padding?: number }} options - Layout.
 */
export function gridForBounds(bounds, { elementSize, padding = DEFAULT_PADDING } = {}) {
    if (!(elementSize > 0 && Number.isFinite(elementSize))) throw elementSizeError(elementSize);
    const h = elementSize;
    const dims = [0, 1, 2].map(a => Math.max(1, Math.ceil((bounds.max[a] - bounds.min[a]) / h - 1e-6)) + 2 * padding);
    const origin = bounds.min.map(v => v - h * padding);
    return { dims, origin, h, occupancy: new Uint8Array(dims[0] * dims[1] * dims[2]) };
}

/** Linear cell index (voxelize.js:76-78). */
export function cellIndex(grid, i, j, k) {
    return i + grid.dims[0] * (j + grid.dims[1] * k);
}

/** Number of occupied cells (voxelize.js:84-90). */
export function countOccupied(grid) {
    let count = 0;
    for (let i = 0; i < grid.occupancy.length; i++) count += grid.occupancy[i];
    return count;
}

/** The addon (or a stand-in passed as options.native) for a call. */
export function nativeOf(options) {
    return options.native || addon();
}

/** Runs a native voxeliser; an elementSize that is no number is reported as given. */
export function withElementSize(options, run) {
    const given = options.elementSize;
    const size = typeof given === 'number' ? given : NaN;
    try {
        return run(size);
    } catch (error) {
        if (typeof given !== 'number' && /elementSize/.test(error.message)) throw elementSizeError(given);
        throw error;
    }
}

function float32(a) {
    return a instanceof Float32Array ? a : Float32Array.from(a);
}

/**
 * voxelizeSdf (voxelize.js:118-148) of an analytic shape from
 * discretize/shapes.mjs (sphere, box, torus, translated up to four times).
 * A shape given only as an sdf closure cannot run on the device.
 * @param {object} shape - The shape; its `device` record names it.
 * @param {{ elementSize: number, padding?: number, device?: number }} options - Layout.
 */
export function voxelizeSdf(shape, options = {}) {
    const d = shape.device;
    if (!d) {
        throw new Error('voxelizeSdf: the device voxeliser takes the analytic shapes of discretize/shapes.mjs, ' +
                        'not an arbitrary sdf');
    }
    const { padding = DEFAULT_PADDING, device = 0 } = options;
    const offsets = new Float64Array(3 * d.offsets.length);
    d.offsets.forEach((o, t) => offsets.set(o, 3 * t));
    return withElementSize(options, size => nativeOf(options).voxelizeShape(
        device, SHAPE_KINDS[d.kind], Float64Array.from(d.size), offsets, size, padding));
}

/**
 * voxelizeMesh (voxelize.js:174-203): two of three axis-parity votes or a
 * surface crossing, then a morphological close.
 * @param {{ positions: Float32Array, indices: Uint32Array }} mesh - Triangles.
 * @param {{ min: number[], max: number[] }} bounds - Bounds of the mesh (meshBounds).
 * @param {{ elementSize: number, padding?: number, close?: boolean, device?: number }} options - Layout.
 */
export function voxelizeMesh(mesh, bounds, options = {}) {
    const { padding = DEFAULT_PADDING, close = true, device = 0 } = options;
    const b = bounds ? Float64Array.from([...bounds.min, ...bounds.max]) : null;
    const indices = mesh.indices instanceof Uint32Array ? mesh.indices : Uint32Array.from(mesh.indices);
    return withElementSize(options, size => nativeOf(options).voxelizeMesh(
        device, float32(mesh.positions), indices, b, size, padding, close));
}

/** meshBounds (body/restGeometry.js:29-40). */
export function meshBounds(positions) {
    const min = [Infinity, Infinity, Infinity];
    const max = [-Infinity, -Infinity, -Infinity];
    for (let i = 0; i < positions.length; i += 3) {
        for (let axis = 0; axis < 3; axis++) {
            const v = positions[i + axis];
            if (v < min[axis]) min[axis] = v;
            if (v > max[axis]) max[axis] = v;
        }
    }
    return { min, max };
}
