// Embedded Gaussians carried by a body's lattice: means by trilinear weights,
// covariances by F Sigma F^T (the maths of packages/visionary-sim/src/sim/
// discretize/lattice/sample.wgsl:40-80 and deformation/covariance.wgsl:30-42),
// run as one device pass over the splat buffers instead of inside a renderer.
import { tileWords } from '../state/particles.mjs';

const SPLAT_TILE = 256;

export class SplatDeformer {
    /**
     * @param {import('../simulator.mjs').Simulator} sim - Owner of the lattice.
     * @param {number} bodyIndex - Which body the splats ride.
     * @param {{ means: Float32Array, quats: Float32Array, scales: Float32Array }} splats -
     *   xyz, xyzw and xyz per splat.
     */
    constructor(sim, bodyIndex, { means, quats, scales }) {
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
        this.output = backend.buffer(9 * 4 * Math.max(1, G));
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
        pass.setupDispatch(count);
        this.pass = pass;
    }

    /** Enqueues the deformation of every splat at the current node positions. */
    deform() { this.backend.dispatch([this.pass], 'splat-deform'); }

    /** @returns {Promise<{ means: Float32Array, cov: Float32Array }>} Planes: 3 of means, 6 of covariance (xx xy xz yy yz zz). */
    async read() {
        const out = new Float32Array(9 * this.count);
        await this.output.read(0, out.byteLength, out);
        return { means: out.subarray(0, 3 * this.count), cov: out.subarray(3 * this.count) };
    }

    destroy() {
        for (const b of [this.input, this.output, this.tiles]) b.destroy();
    }
}
