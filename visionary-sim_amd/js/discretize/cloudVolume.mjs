// voxelizeCloud (packages/visionary-sim/src/sim/discretize/lattice/
// cloudVolume.js:95-121) on the HIP device: the Gaussians stamped as solid
// ellipsoids, then sealed, flooded from the faces and eroded back.
import { DEFAULT_PADDING, nativeOf, withElementSize } from './voxelize.mjs';

/** Standard deviations of a Gaussian that count as material (cloudVolume.js:70). */
export const SIGMA = 2;
/** Cells the shell is thickened by before the fill (cloudVolume.js:83). */
export const SEAL = 2;

/**
 * @param {{ count: number, positions: Float32Array, rotations: Float32Array, scales: Float32Array }} cloud -
 *   Gaussians in world space, xyzw unit rotations, scales in metres.
 * @param {{ elementSize: number, padding?: number, sigma?: number, seal?: number, fill?: boolean,
 *   device?: number }} options - Layout and method.
 */
export function voxelizeCloud(cloud, options = {}) {
    const { sigma = SIGMA, seal = SEAL, fill = true, padding = DEFAULT_PADDING, device = 0 } = options;
    if (cloud.count === 0) throw new Error('voxelizeCloud: a cloud with no Gaussians has no volume');
    const n = cloud.count;
    const f32 = (a, w) => (a instanceof Float32Array ? a : Float32Array.from(a)).subarray(0, w * n);
    return withElementSize(options, size => nativeOf(options).voxelizeCloud(
        device, f32(cloud.positions, 3), f32(cloud.rotations, 4), f32(cloud.scales, 3), size, padding, sigma, seal,
        fill));
}
