// Grabbing, answered on the device: which node a ray picks
// (nearestNodeToRay, packages/visionary-sim/src/sim/pick.js:41-68) and which
// nodes the hand then holds (withinRadius, grip.js:99-138), as one launch
// each over the live position buffer. A grab reads back the anchor, its
// position and the held set -- a few hundred bytes -- instead of every node
// position (simulator.js:383-425 reads them all).
//
// The arithmetic is the reference's: f64 over the f32 positions, so the held
// set and its offsets are the ones the JS functions would compute.

/** Default hand radius, metres (grip.js:155-172). */
export const DEFAULT_GRIPPED_PART_RADIUS = 0.03;

/** Upper bound on the nodes a `radius` ball can hold at `spacing` (grip.js:174-199). */
export function worstCaseHeldNodes(radius, spacing) {
    return Math.floor(((2 * radius) / spacing + 1) ** 3);
}

/** Which body owns a node, given where each body's nodes start (pick.js:77-83). */
export function bodyOfNode(nodeOffsets, node) {
    let body = 0;
    while (body + 1 < nodeOffsets.length && node >= nodeOffsets[body + 1]) body++;
    return body;
}

/** Strategies a grab can use to turn an anchor into a held set (grip.js:141). */
export const GRIP_STRATEGIES = ['singleNode', 'withinRadius'];

const PICK_GROUPS = 64;
const GRIP_GROUPS = 32;
const QUERY_BYTES = 72;  // vs_grab_params

export class GrabQueries {
    /**
     * @param {import('./compute/backend.mjs').ComputeBackend} backend
     * @param {import('./state/particles.mjs').Particles} particles
     * @param {number} capacity - Most nodes a grab may hold (the drive's capacity).
     */
    constructor(backend, particles, capacity) {
        this.backend = backend;
        this.particles = particles;
        this.capacity = capacity;
        const N = Math.max(1, particles.nodeCount);
        this.query = backend.buffer(QUERY_BYTES);
        this.pickOut = backend.buffer(16 + 16 * PICK_GROUPS);
        this.candidates = backend.buffer(16 * N);
        this.gripOut = backend.buffer(32 + 4 * (capacity + (capacity & 1)) + 24 * capacity);
        this.pickPass = backend.pass(backend.kernel('sim-pick-ray'), 'sim-pick-ray');
        this.gripPass = backend.pass(backend.kernel('sim-grip-radius'), 'sim-grip-radius');
        for (const pass of [this.pickPass, this.gripPass]) {
            pass.set('nodePos', particles.nodePos);
            pass.set('grabMass', particles.nodeGrabMass);
            pass.set('query', this.query);
            pass.set('plane', particles.nodeCount);
        }
        this.pickPass.set('pickOut', this.pickOut);
        this.pickPass.setupDispatch(PICK_GROUPS);
        this.gripPass.set('candidates', this.candidates);
        this.gripPass.set('gripOut', this.gripOut);
        this.gripPass.set('items', N);
        this.gripPass.setupDispatch(GRIP_GROUPS);
    }

    writeQuery({ origin = [0, 0, 0], direction = [0, 0, 0], radius = 0, anchor = 0, begin = 0, end = 0 }) {
        const bytes = new ArrayBuffer(QUERY_BYTES);
        const f = new Float64Array(bytes, 0, 7);
        f.set(origin, 0);
        f.set(direction, 3);
        f[6] = radius;
        const u = new Uint32Array(bytes, 56, 4);
        u[0] = anchor;
        u[1] = begin;
        u[2] = end;
        this.query.write(0, new Uint8Array(bytes));
    }

    /**
     * @returns {Promise<{ node: number, depth: number }>} The picked node (-1 for a miss).
     */
    async pick(origin, direction, radius) {
        this.writeQuery({ origin, direction, radius });
        this.backend.dispatch([this.pickPass], 'grab-pick');
        const out = new ArrayBuffer(16);
        await this.pickOut.read(0, 16, new Uint8Array(out));
        return { node: new Int32Array(out, 4, 1)[0], depth: new Float64Array(out, 8, 1)[0] };
    }

    /**
     * withinRadius on the device.
     * @returns {Promise<{ nodes: number[], offsets: number[][], found: number }>}
     */
    async grip(anchor, { radius, nodeOffsets, limit = this.capacity }) {
        const body = bodyOfNode(nodeOffsets, anchor);
        const begin = nodeOffsets[body];
        const end = body + 1 < nodeOffsets.length ? nodeOffsets[body + 1] : this.particles.nodeCount;
        const cap = Math.min(limit, this.capacity);
        this.writeQuery({ radius, anchor, begin, end });
        this.gripPass.set('pairCap', cap);
        this.backend.dispatch([this.gripPass], 'grab-grip');
        const head = new Uint32Array(8);
        await this.gripOut.read(0, 32, head);
        const [, , found, kept, status] = head;
        if (status === 2) throw new Error(`[grip] ${found} nodes within ${radius} m exceed what one grab can sort`);
        if (status === 3) throw new Error(`[grip] anchor ${anchor} outside its body's node range`);
        const nodes = new Uint32Array(Math.max(1, kept));
        const offs = new Float64Array(Math.max(1, 3 * kept));
        if (kept) {
            await this.gripOut.read(32, 4 * kept, nodes);
            await this.gripOut.read(32 + 4 * (cap + (cap & 1)), 24 * kept, offs);
        }
        const offsets = [];
        for (let i = 0; i < kept; i++) offsets.push([offs[3 * i], offs[3 * i + 1], offs[3 * i + 2]]);
        return { nodes: Array.from(nodes.subarray(0, kept)), offsets, found };
    }

    /** The live position of one node (three 4-byte reads). */
    async position(node) {
        const N = this.particles.nodeCount;
        const v = new Float32Array(1);
        const p = [];
        for (let a = 0; a < 3; a++) {
            await this.particles.nodePos.read(4 * (a * N + node), 4, v);
            p.push(v[0]);
        }
        return p;
    }

    destroy() {
        for (const b of [this.query, this.pickOut, this.candidates, this.gripOut]) b.destroy();
    }
}
