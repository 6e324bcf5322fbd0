// Embedded Gaussians carried by a body's lattice: means by trilinear weights,
// covariances by F Sigma F^T (the maths of packages/visionary-sim/src/sim/
// discretize/lattice/sample.wgsl:40-80 and deformation/covariance.wgsl:30-42),
// run as one device pass over the splat buffers instead of inside a renderer.
import { tileWords } from '../state/particles.mjs';

const SPLAT_TILE = 256;
// Outputs of the pass (VS_SPLAT_* of include/vsim.h): the covariance, and the
// engine's splat format -- a unit rotation and three axis lengths, the
// covariance taken apart by four sweeps of cyclic Jacobi
// (apps/simulator/src/render/ellipsoid.wgsl:32-91).
export const SPLAT_COV = 1;
export const SPLAT_AXES = 2;

export class SplatDeformer {
    /**
     * @param {import('../simulator.mjs').Simulator} sim - Owner of the lattice.
     * @param {number} bodyIndex - Which body the splats ride.
     * @param {{ means: Float32Array, quats: Float32Array, scales: Float32Array }} splats -
     *   xyz, xyzw and xyz per splat.
     * @param {{ axes?: boolean }} [options] - `axes`: also write the engine format.
     */
    constructor(sim, bodyIndex, { means, quats, scales }, { axes = false } = {}) {
        const backend = sim.backend;
        const G = means.length / 3;
        this.count = G;
        this.backend = backend;
        const input = new Float32Array(10 * Math.max(1, G));
        input.set(quats.subarray(0, 4 * G), 0);
        for (let i = 0; i < G; i++) {
            for (let a = 0; a < 3; a++) {
                input[(4 + a) * G + i] = means[3 * i + a];
                input[(7 + a) * G + i] = scales[3 * i + a];
            }
        }
        this.input = backend.buffer(input.byteLength);
        this.input.write(0, input);
        this.axes = axes;
        this.output = backend.buffer((axes ? 16 : 9) * 4 * Math.max(1, G));
        const { words, count } = tileWords([{ begin: 0, count: G, body: bodyIndex }], SPLAT_TILE);
        this.tiles = backend.buffer(words.byteLength);
        this.tiles.write(0, words);
        const pass = backend.pass(backend.kernel('sim-splat-deform'), 'sim-splat-deform');
        pass.set('nodePos', sim.particles.nodePos);
        pass.set('bodies', sim.particles.bodyTable);
        pass.set('tiles', this.tiles);
        pass.set('splatIn', this.input);
        pass.set('splatOut', this.output);
        pass.set('items', G);
        pass.set('plane', sim.particles.nodeCount);
        pass.set('block', SPLAT_TILE);
        pass.set('outputs', axes ? SPLAT_COV | SPLAT_AXES : SPLAT_COV);
        pass.setupDispatch(count);
        this.pass = pass;
    }

    /** Enqueues the deformation of every splat at the current node positions. */
    deform() { this.backend.dispatch([this.pass], 'splat-deform'); }

    /** @returns {Promise<{ means: Float32Array, cov: Float32Array, rotation?: Float32Array, scale?: Float32Array }>}
     *  Planes: 3 of means, 6 of covariance (xx xy xz yy yz zz); with `axes`, 4 of rotation (xyzw) and 3 of axis lengths. */
    async read() {
        const G = this.count;
        const out = new Float32Array((this.axes ? 16 : 9) * G);
        await this.output.read(0, out.byteLength, out);
        const result = { means: out.subarray(0, 3 * G), cov: out.subarray(3 * G, 9 * G) };
        if (this.axes) {
            result.rotation = out.subarray(9 * G, 13 * G);  // planes x, y, z, w
            result.scale = out.subarray(13 * G, 16 * G);
        }
        return result;
    }

    destroy() {
        for (const b of [this.input, this.output, this.tiles]) b.destroy();
    }
}
