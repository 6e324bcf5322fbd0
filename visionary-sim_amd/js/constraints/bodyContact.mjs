// Contact between bodies: bin, canonical sort, Jacobi gather, settle (the
// role of packages/visionary-sim/src/sim/constraints/bodyContact.js:50-214).
// Surface records are nine planes of surfaceCount words: node, body, radius,
// bucket, arrival rank, gathered correction x y z and hit count.
import { tileWords } from '../state/particles.mjs';
import { bindAccepted } from './elasticCells.mjs';

export const BUCKET_SLOTS = 32;
const GAIN = 1;
const SURFACE_TILE = 256;

/** broadphaseGrid (broadphase.js:46-58): one contact diameter per bucket plus a margin ring. */
export function broadphaseGrid(bounds, bucketSize) {
    const edge = bucketSize > 0 ? bucketSize : Math.max(...bounds.min.map((min, a) => bounds.max[a] - min));
    const dims = bounds.min.map((min, a) => Math.max(1, Math.ceil((bounds.max[a] - min) / edge)) + 2);
    return { origin: bounds.min.map(min => min - edge), bucketSize: edge, dims, bucketCount: dims[0] * dims[1] * dims[2] };
}

/** Hard bound on one body's nodes per bucket (broadphase.js:82-85). */
export function worstCaseBucketOccupancy(bucketSize, spacing) {
    return Math.pow(Math.ceil((Math.sqrt(3) * bucketSize) / spacing), 3);
}

export class BodyContact {
    constructor(backend, particles, proxies, bounds, options = {}) {
        this.particles = particles;
        this.enabled = true;
        const records = [];
        let widest = 0;
        let finest = Infinity;
        proxies.forEach((proxy, b) => {
            const { nodes, radii } = proxy.contactSurface();
            const offset = particles.offsetOf(b);
            for (let s = 0; s < nodes.length; s++) {
                widest = Math.max(widest, radii[s]);
                records.push({ node: offset + nodes[s], body: b, radius: radii[s] });
            }
            finest = Math.min(finest, proxy.spacing.min);
        });
        this.surfaceCount = records.length;
        this.grid = broadphaseGrid(bounds, options.bucketSize || widest * 2);
        if (Number.isFinite(finest) && worstCaseBucketOccupancy(this.grid.bucketSize, finest) > BUCKET_SLOTS) {
            console.warn(`[contact] a bucket of ${this.grid.bucketSize} m can hold more than ${BUCKET_SLOTS} nodes of the finest body`);
        }
        const S = Math.max(1, records.length);
        const planes = new ArrayBuffer(9 * 4 * S);
        const u32 = new Uint32Array(planes);
        const f32 = new Float32Array(planes);
        records.forEach((r, i) => {
            u32[i] = r.node;
            u32[S + i] = r.body;
            f32[2 * S + i] = r.radius;
        });
        this.surface = backend.buffer(planes.byteLength);
        this.surface.write(0, new Uint8Array(planes));
        this.contactGrid = backend.buffer(4 * this.grid.bucketCount * (1 + BUCKET_SLOTS));
        const { words, count } = tileWords([{ begin: 0, count: records.length, body: 0 }], SURFACE_TILE);
        this.tiles = backend.buffer(words.byteLength);
        this.tiles.write(0, words);
        const u = particles.uniforms;
        u.set('gridOrigin', this.grid.origin);
        u.set('gridDims', this.grid.dims);
        u.set('bucketSize', this.grid.bucketSize);
        u.set('bucketCount', this.grid.bucketCount);
        u.set('gridBase', 0);
        u.set('gain', GAIN);
        u.set('surfBase', 0);
        u.set('surfaceCount', records.length);
        u.set('bucketSlots', BUCKET_SLOTS);
        this.pairLog = null;
        const make = name => {
            const pass = backend.pass(backend.kernel(`sim-contact-${name}`), `sim-contact-${name}`);
            bindAccepted(pass, particles.bindings);
            pass.set('surface', this.surface);
            pass.set('contactGrid', this.contactGrid);
            pass.set('items', S);
            pass.set('buckets', this.grid.bucketCount);
            pass.set('tiles', this.tiles);
            pass.set('block', SURFACE_TILE);
            pass.setupDispatch(count);
            return pass;
        };
        this.dispatches = records.length === 0 ? [] : ['bin', 'sort', 'collide', 'settle'].map(make);
    }

    /** Records the pairs the collide pass accepts, `cap` per record. */
    logPairs(backend, cap) {
        const S = Math.max(1, this.surfaceCount);
        this.pairLog = backend.buffer(4 * S * (1 + cap));
        const collide = this.dispatches[2];
        collide.set('pairLog', this.pairLog);
        collide.set('pairCap', cap);
    }

    get passes() { return this.enabled ? this.dispatches : []; }

    setSubstepDt() {}

    destroy() {
        for (const b of [this.surface, this.contactGrid, this.tiles]) b.destroy();
        if (this.pairLog) this.pairLog.destroy();
    }
}
