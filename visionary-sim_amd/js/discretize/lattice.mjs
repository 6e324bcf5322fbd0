// The voxel lattice: dense node grid, corner-lumped masses, cells grouped by
// parity colour (semantics of packages/visionary-sim/src/sim/discretize/
// lattice/lattice.js:44-239 and voxelLattice.js:17-112). Host setup only; it
// fixes the order of every GPU input.

import { meshBounds, voxelizeMesh, voxelizeSdf } from './voxelize.mjs';
import { voxelizeCloud } from './cloudVolume.mjs';

export const COLOR_COUNT = 8;
/** Contact sphere radius as a fraction of the cell (voxelLattice.js:17). */
export const CONTACT_RADIUS_CELLS = 0.5;

export function cellColor(i, j, k) {
    return (i & 1) | ((j & 1) << 1) | ((k & 1) << 2);
}

export function nodeIndex(nodeDims, i, j, k) {
    return i + nodeDims[0] * (j + nodeDims[1] * k);
}

/**
 * @param {{ dims: number[], origin: number[], h: number, occupancy: Uint8Array }} grid -
 *   Occupancy, linear index i + dims[0] * (j + dims[1] * k), nonzero = material.
 * @param {{ density: number }} options - Material density.
 */
export function buildLattice(grid, { density }) {
    const { dims: cellDims, origin, h, occupancy } = grid;
    const nodeDims = [cellDims[0] + 1, cellDims[1] + 1, cellDims[2] + 1];
    const nodeCount = nodeDims[0] * nodeDims[1] * nodeDims[2];
    const restPositions = new Float32Array(nodeCount * 3);
    for (let k = 0; k < nodeDims[2]; k++) {
        for (let j = 0; j < nodeDims[1]; j++) {
            for (let i = 0; i < nodeDims[0]; i++) {
                const n = nodeIndex(nodeDims, i, j, k) * 3;
                restPositions[n] = origin[0] + i * h;
                restPositions[n + 1] = origin[1] + j * h;
                restPositions[n + 2] = origin[2] + k * h;
            }
        }
    }
    const cornerMass = (density * h * h * h) / 8;
    const mass = new Float32Array(nodeCount);
    const byColor = [];
    for (let c = 0; c < COLOR_COUNT; c++) byColor.push([]);
    let occupied = 0;
    for (let k = 0; k < cellDims[2]; k++) {
        for (let j = 0; j < cellDims[1]; j++) {
            for (let i = 0; i < cellDims[0]; i++) {
                const cell = i + cellDims[0] * (j + cellDims[1] * k);
                if (!occupancy[cell]) continue;
                occupied++;
                byColor[cellColor(i, j, k)].push(cell);
                for (let c = 0; c < 8; c++) {
                    mass[nodeIndex(nodeDims, i + (c & 1), j + ((c >> 1) & 1), k + ((c >> 2) & 1))] += cornerMass;
                }
            }
        }
    }
    const cells = new Uint32Array(occupied);
    const colorRanges = [];
    let start = 0;
    for (const group of byColor) {
        cells.set(group, start);
        colorRanges.push({ start, count: group.length });
        start += group.length;
    }
    const invMass = new Float32Array(nodeCount);
    for (let n = 0; n < nodeCount; n++) invMass[n] = mass[n] > 0 ? 1 / mass[n] : 0;
    return {
        origin: origin.slice(), h, cellDims: cellDims.slice(), nodeDims, nodeCount, restPositions, invMass,
        cells, colorRanges, mass: occupied * density * h * h * h, volume: occupied * h * h * h
    };
}

/** Nodes with 1..7 of their 8 incident cells occupied, ascending. */
export function surfaceNodes({ cellDims, nodeDims, cells }) {
    const occupied = new Uint8Array(cellDims[0] * cellDims[1] * cellDims[2]);
    for (const cell of cells) occupied[cell] = 1;
    const found = [];
    for (let k = 0; k < nodeDims[2]; k++) {
        for (let j = 0; j < nodeDims[1]; j++) {
            for (let i = 0; i < nodeDims[0]; i++) {
                let filled = 0;
                for (let c = 0; c < 8; c++) {
                    const ci = i + (c & 1) - 1;
                    const cj = j + ((c >> 1) & 1) - 1;
                    const ck = k + ((c >> 2) & 1) - 1;
                    if (ci >= 0 && cj >= 0 && ck >= 0 && ci < cellDims[0] && cj < cellDims[1] && ck < cellDims[2]) {
                        filled += occupied[ci + cellDims[0] * (cj + cellDims[1] * ck)];
                    }
                }
                if (filled > 0 && filled < 8) found.push(nodeIndex(nodeDims, i, j, k));
            }
        }
    }
    return Uint32Array.from(found);
}

/** A full block of `cells` = [x, y, z] cells. */
export function blockGrid(cells, origin, h) {
    const occupancy = new Uint8Array(cells[0] * cells[1] * cells[2]).fill(1);
    return { dims: cells.slice(), origin: origin.slice(), h, occupancy };
}

/**
 * The lattice as the simulator's proxy geometry: degrees of freedom, spacing,
 * CFL length and contact surface (voxelLattice.js:278-316).
 */
export class VoxelLattice {
    constructor(lattice) {
        this.lattice = lattice;
        this.dof = { restPositions: lattice.restPositions, invMass: lattice.invMass, nodeCount: lattice.nodeCount };
        this.spacing = { min: lattice.h, max: lattice.h };
        this.cflLength = lattice.h;
        this.surface = null;
    }

    contactSurface() {
        if (!this.surface) {
            const nodes = surfaceNodes(this.lattice);
            const radii = new Float32Array(nodes.length).fill(this.lattice.h * CONTACT_RADIUS_CELLS);
            this.surface = { nodes, radii };
        }
        return this.surface;
    }

    setDensity(previous, density) {
        const ratio = previous / density;
        const { invMass, nodeCount } = this.lattice;
        for (let n = 0; n < nodeCount; n++) invMass[n] *= ratio;
        this.lattice.mass = this.lattice.volume * density;
    }
}

/**
 * How each kind of rest geometry becomes an occupancy grid (voxelLattice.js:
 * 208-212): meshes, analytic shapes and Gaussian clouds are voxelised on the
 * backend's device; an occupancy grid or a block is taken as given.
 */
const VOXELIZERS = {
    mesh: (source, options) => voxelizeMesh(source.mesh, meshBounds(source.mesh.positions), options),
    shape: (source, options) => voxelizeSdf(source.shape, options),
    cloud: (source, options) => voxelizeCloud(source.cloud, options),
    lattice: source => source.grid,
    block: source => blockGrid(source.cells, source.origin, source.h)
};

/**
 * proxyFor (voxelLattice.js:218-227): the body's lattice, cells
 * `body.elementSize` across.
 * @param {object} body - A SoftBody.
 * @param {object} [backend] - The ComputeBackend whose device and addon voxelise.
 */
export function proxyFor(body, backend = null) {
    const { source, material, elementSize } = body;
    const voxelize = VOXELIZERS[source.kind];
    if (!voxelize) throw new Error(`no way to discretise rest geometry of kind "${source.kind}"`);
    const options = { elementSize };
    if (backend) Object.assign(options, { device: backend.device, native: backend.native });
    return new VoxelLattice(buildLattice(voxelize(source, options), { density: material.density }));
}
