// Every degree of freedom of the scene as structure-of-arrays device buffers
// (the role of packages/visionary-sim/src/sim/state/particles.js:55-256):
// nodePos / nodePrev / nodeVel hold three planes (x, y, z) of nodeCount
// floats, nodeInvMass one plane. The per-body record (vs_body) carries what
// the node passes read (offset, count, friction, damping factor) and what the
// elastic pass reads (written by ElasticCells). Composition of the GPU inverse
// mass from material truth and the static / kinematic prescriptions follows
// particles.js:206-215.
import { BODY, BODY_STRIDE, TILE_WORDS } from '../native.mjs';

/** Nodes per node-pass tile (one workgroup). */
export const NODE_TILE = 256;

/** exp(-damping * dt) as finalize.wgsl:72 forms it: the f32 product of the two uniforms, exponentiated once. */
export function dampFactor(damping, dt) {
    return Math.fround(Math.exp(Math.fround(-(Math.fround(damping) * Math.fround(dt)))));
}

/** Tile buffer contents: header {count, 0, 0, 0} then {begin, count, body, scene} records. */
export function tileWords(ranges, tile) {
    const list = [];
    for (const { begin, count, body } of ranges) {
        for (let o = 0; o < count; o += tile) list.push([begin + o, Math.min(tile, count - o), body, 0]);
    }
    const words = new Uint32Array(TILE_WORDS * (1 + list.length));
    words[0] = list.length;
    list.forEach((t, i) => words.set(t, TILE_WORDS * (1 + i)));
    return { words, count: list.length };
}

export class Particles {
    constructor(backend, bodies, uniforms) {
        this.backend = backend;
        this.bodies = bodies;
        this.uniforms = uniforms;
        let total = 0;
        this.offsets = bodies.map(({ nodeCount }) => {
            const offset = total;
            total += nodeCount;
            return offset;
        });
        this.nodeCount = total;
        const N = total;
        this.restPlanes = new Float32Array(3 * N);
        this.invMass = new Float32Array(N);
        for (let b = 0; b < bodies.length; b++) {
            const { restPositions, nodeCount } = bodies[b];
            const offset = this.offsets[b];
            for (let n = 0; n < nodeCount; n++) {
                for (let a = 0; a < 3; a++) this.restPlanes[a * N + offset + n] = restPositions[n * 3 + a];
            }
            this.invMass.set(bodies[b].invMass, offset);
        }
        this.staticPrescribed = new Set();
        bodies.forEach(({ staticNodes = [] }, b) => staticNodes.forEach(n => this.staticPrescribed.add(this.offsets[b] + n)));
        this.prescribed = new Set();
        this.gpuInvMass = new Float32Array(this.invMass);

        this.nodePos = backend.buffer(12 * N);
        this.nodePrev = backend.buffer(12 * N);
        this.nodeVel = backend.buffer(12 * N);
        this.nodeInvMass = backend.buffer(4 * N);
        // What a grab may take (isGrabbableNode, pick.js:14-24): the material
        // inverse mass with static nodes zeroed. Held nodes stay grabbable.
        this.nodeGrabMass = backend.buffer(4 * N);
        this.bodyWords = new ArrayBuffer(BODY_STRIDE * bodies.length);
        this.bodyF32 = new Float32Array(this.bodyWords);
        this.bodyU32 = new Uint32Array(this.bodyWords);
        this.bodyTable = backend.buffer(BODY_STRIDE * bodies.length);
        this.params = bodies.map(() => ({ damping: 0, friction: 0 }));
        this.dt = 0;
        for (let b = 0; b < bodies.length; b++) {
            const base = b * (BODY_STRIDE / 4);
            this.bodyU32[base + BODY.nodeOffset] = this.offsets[b];
            this.bodyU32[base + BODY.nodeCount] = bodies[b].nodeCount;
            this.bodyU32[base + BODY.scene] = 0;
        }
        const { words, count } = tileWords(bodies.map((b, i) => ({ begin: this.offsets[i], count: b.nodeCount, body: i })),
            NODE_TILE);
        this.tiles = backend.buffer(words.byteLength);
        this.tiles.write(0, words);
        this.tileCount = count;
        this.reset();
        this.uploadGpuMasses();
        this.uploadGrabMass();
        this.uploadParams();
    }

    offsetOf(bodyIndex) { return this.offsets[bodyIndex]; }

    /** `{damping, friction}` of one body (particles.js:139-142). */
    setParams(bodyIndex, values) {
        Object.assign(this.params[bodyIndex], values);
        this.uploadParams();
    }

    /** The damping factor depends on the substep; recomputed when it changes. */
    setSubstepDt(dt) {
        if (dt === this.dt) return;
        this.dt = dt;
        this.uploadParams();
    }

    /** Rewrites the body records (both halves) from the host copy. */
    uploadParams() {
        for (let b = 0; b < this.bodies.length; b++) {
            const base = b * (BODY_STRIDE / 4);
            this.bodyF32[base + BODY.friction] = this.params[b].friction;
            this.bodyF32[base + BODY.damp] = this.dt > 0 ? dampFactor(this.params[b].damping, this.dt) : 1;
        }
        this.bodyTable.write(0, new Uint8Array(this.bodyWords));
    }

    setPrescribed(nodes) {
        this.prescribed = new Set(nodes);
        this.uploadGpuMasses();
    }

    clearPrescribed() {
        this.prescribed = new Set();
        this.uploadGpuMasses();
    }

    uploadMasses() {
        for (let b = 0; b < this.bodies.length; b++) this.invMass.set(this.bodies[b].invMass, this.offsets[b]);
        this.uploadGpuMasses();
        this.uploadGrabMass();
    }

    uploadGrabMass() {
        const grab = new Float32Array(this.invMass);
        for (const n of this.staticPrescribed) grab[n] = 0;
        this.nodeGrabMass.write(0, grab);
    }

    uploadGpuMasses() {
        this.gpuInvMass.set(this.invMass);
        for (const n of this.staticPrescribed) this.gpuInvMass[n] = 0;
        for (const n of this.prescribed) this.gpuInvMass[n] = 0;
        this.nodeInvMass.write(0, this.gpuInvMass);
    }

    /** Positions, 4 floats per node (xyzw, w = 0), like particles.js:227-231. */
    async readPositions() {
        const N = this.nodeCount;
        const planes = new Float32Array(3 * N);
        await this.nodePos.read(0, planes.byteLength, planes);
        const out = new Float32Array(4 * N);
        for (let n = 0; n < N; n++) {
            out[4 * n] = planes[n];
            out[4 * n + 1] = planes[N + n];
            out[4 * n + 2] = planes[2 * N + n];
        }
        return out;
    }

    /** Velocity planes (x, y, z), for diagnostics. */
    async readVelocityPlanes() {
        const planes = new Float32Array(3 * this.nodeCount);
        await this.nodeVel.read(0, planes.byteLength, planes);
        return planes;
    }

    get bindings() {
        return {
            nodePos: this.nodePos, nodePrev: this.nodePrev, nodeVel: this.nodeVel, nodeInvMass: this.nodeInvMass,
            bodies: this.bodyTable, params: this.uniforms, plane: this.nodeCount
        };
    }

    reset() {
        this.nodePos.write(0, this.restPlanes);
        this.nodePrev.write(0, this.restPlanes);
        const vel = new Float32Array(3 * this.nodeCount);
        this.bodies.forEach((b, i) => {
            const v = b.velocity || [0, 0, 0];
            for (let a = 0; a < 3; a++) vel.fill(v[a], a * this.nodeCount + this.offsets[i], a * this.nodeCount + this.offsets[i] + b.nodeCount);
        });
        this.nodeVel.write(0, vel);
    }

    destroy() {
        for (const b of [this.nodePos, this.nodePrev, this.nodeVel, this.nodeInvMass, this.nodeGrabMass, this.bodyTable,
            this.tiles]) b.destroy();
    }
}
