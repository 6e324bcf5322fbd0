// The HIP ComputeBackend: the same surface as the kernel package's
// private compute boundary (packages/visionary-sim/src/sim/compute/
// backend.js:114-163) -- buffer / kernel / pass / dispatch / destroy -- over
// the Node-API addon. Kernels are keyed by the reference's pass labels; the
// WGSL text argument is accepted and ignored. A pass's uniforms live in a
// shared device record (vs_scene_params) uploaded once per dispatch when it
// changed, as backend.js:142-148 uploads each pass's uniform buffer.
import { addon, SLOT, LAUNCH_BYTES, SCENE, SCENE_BYTES } from '../native.mjs';

const roundUp = (n, a) => Math.ceil(n / a) * a;

const NODE = ['nodePos', 'nodePrev', 'nodeInvMass', 'params', 'tiles', 'plane', 'block'];
/** Parameters each kernel takes: buffer slots by name and launch scalars. */
const KERNEL_PARAMS = {
    'sim-predict': [...NODE, 'nodeVel'],
    'sim-finalize': [...NODE, 'nodeVel', 'bodies'],
    'sim-elastic-cells': [...NODE, 'bodies', 'cells', 'limitOnly'],
    'sim-contact-bin': [...NODE, 'surface', 'contactGrid', 'items', 'buckets'],
    'sim-contact-sort': [...NODE, 'surface', 'contactGrid', 'items', 'buckets'],
    'sim-contact-collide': [...NODE, 'bodies', 'surface', 'contactGrid', 'items', 'buckets', 'pairLog', 'pairCap'],
    'sim-contact-settle': [...NODE, 'surface', 'contactGrid', 'items', 'buckets'],
    'sim-drive-hard': [...NODE, 'nodeVel', 'driven', 'items'],
    'sim-drive-compliant': [...NODE, 'nodeVel', 'driven', 'items'],
    'sim-splat-deform': ['nodePos', 'bodies', 'tiles', 'splatIn', 'splatOut', 'items', 'plane', 'block', 'outputs'],
    'sim-pick-ray': ['nodePos', 'grabMass', 'query', 'pickOut', 'plane', 'block'],
    'sim-grip-radius': ['nodePos', 'grabMass', 'query', 'candidates', 'gripOut', 'plane', 'items', 'pairCap', 'block']
};
const OPTIONAL = new Set(['block', 'pairLog', 'pairCap', 'limitOnly', 'nodeInvMass', 'nodePrev', 'params', 'outputs']);
const BUFFER_SLOT = {
    nodePos: SLOT.nodePos, nodePrev: SLOT.nodePrev, nodeVel: SLOT.nodeVel, nodeInvMass: SLOT.nodeInvMass,
    bodies: SLOT.bodies, params: SLOT.params, tiles: SLOT.tiles, cells: SLOT.items, surface: SLOT.items,
    driven: SLOT.items, splatIn: SLOT.items, contactGrid: SLOT.contactGrid, splatOut: SLOT.out, pairLog: SLOT.out,
    grabMass: SLOT.nodeInvMass, query: SLOT.params, candidates: SLOT.items, pickOut: SLOT.out, gripOut: SLOT.out
};

class Allocation {
    constructor(backend, size) {
        this.backend = backend;
        this.size = size;
        this.native = backend.native.alloc(backend.handle, Math.max(4, roundUp(size, 4)));
        backend.allocations.add(this);
    }

    write(offset, data) {
        if (data.byteLength) this.backend.native.write(this.backend.handle, this.native, offset, data);
    }

    async read(offset, size, target) {
        if (!size) return target;
        this.backend.native.read(this.backend.handle, this.native, offset,
            new Uint8Array(target.buffer, target.byteOffset, size));
        return target;
    }

    destroy() {
        if (!this.backend.allocations.delete(this)) return;
        this.backend.native.free(this.backend.handle, this.native);
    }
}

/** A device vs_scene_params record with named fields (the passes' uniforms). */
export class Uniforms {
    constructor(backend) {
        this.buffer = backend.buffer(SCENE_BYTES);
        this.words = new ArrayBuffer(SCENE_BYTES);
        this.f32 = new Float32Array(this.words);
        this.u32 = new Uint32Array(this.words);
        this.dirty = true;
        this.set('bucketSlots', 32);
        this.set('gain', 1);
        this.set('responseTime', 0.05);
        backend.uniformSets.add(this);
    }

    has(name) { return Object.prototype.hasOwnProperty.call(SCENE, name); }

    set(name, value) {
        const field = SCENE[name];
        if (!field) throw new Error(`uniforms: unknown parameter ${name}`);
        const [at, type, length] = field;
        const values = length === 1 ? [value] : value;
        if (!values || values.length !== length || values.some(v => !Number.isFinite(v))) {
            throw new Error(`uniforms: invalid parameter ${name}`);
        }
        const view = type === 'f32' ? this.f32 : this.u32;
        for (let i = 0; i < length; i++) view[at + i] = values[i];
        this.dirty = true;
    }

    upload() {
        if (!this.dirty) return;
        this.buffer.write(0, new Uint8Array(this.words));
        this.dirty = false;
    }
}

class Pass {
    constructor(backend, kernel, name) {
        this.backend = backend;
        this.kernel = kernel;
        this.name = name;
        this.values = new Map();
        this.groups = [1, 1, 1];
    }

    set(name, value) {
        if (!this.kernel.params.has(name)) {
            // Uniform fields go to the shared record the pass is bound to.
            const uniforms = this.values.get('params');
            if (uniforms instanceof Uniforms && uniforms.has(name)) {
                uniforms.set(name, value);
                return;
            }
            throw new Error(`${this.name}: unknown parameter ${name}`);
        }
        this.values.set(name, value);
    }

    setupDispatch(x, y = 1, z = 1) {
        const limit = 0x7fffffff;
        if ([x, y, z].some(n => !Number.isInteger(n) || n < 0 || n > limit) || y !== 1 || z !== 1) {
            throw new Error(`${this.name}: invalid dispatch ${[x, y, z]}; device limit is ${limit}`);
        }
        this.groups = [x, y, z];
    }

    encode(view, at) {
        const v = name => this.values.get(name);
        for (const name of this.kernel.params) {
            if (!OPTIONAL.has(name) && v(name) === undefined) throw new Error(`${this.name}: missing parameter ${name}`);
        }
        view.setInt32(at, this.kernel.id, true);
        view.setUint32(at + 4, this.groups[0], true);
        view.setUint32(at + 8, v('block') || 64, true);
        view.setUint32(at + 12, v('plane') || 0, true);
        view.setUint32(at + 16, v('items') || 0, true);
        // flags: the elastic limit-only sweep, or the splat outputs (VS_SPLAT_*).
        view.setUint32(at + 20, v('limitOnly') ? 1 : (v('outputs') || 0), true);
        view.setUint32(at + 24, v('buckets') || 0, true);
        view.setUint32(at + 28, v('pairCap') || 0, true);
        for (const name of this.kernel.params) {
            const slot = BUFFER_SLOT[name];
            if (slot === undefined) continue;
            let buffer = v(name);
            if (buffer === undefined) continue;
            if (buffer instanceof Uniforms) buffer = buffer.buffer;
            if (!(buffer instanceof Allocation) || buffer.backend !== this.backend) {
                throw new Error(`${this.name}: invalid buffer ${name}`);
            }
            view.setBigUint64(at + 32 + 8 * slot, buffer.native, true);
        }
    }

    destroy() { this.values.clear(); }
}

export class ComputeBackend {
    /**
     * @param {number|object} device - HIP device index (or an addon-shaped
     *   object standing in for the native module, as tests do).
     * @param {object} [native] - The addon; defaults to the built vsim.node.
     */
    constructor(device = 0, native = null) {
        this.native = native || addon();
        this.handle = this.native.create(device);
        /** HIP device index (the voxelisers run there too). */
        this.device = typeof device === 'number' ? device : 0;
        this.allocations = new Set();
        this.uniformSets = new Set();
        this.destroyed = false;
    }

    buffer(size) {
        if (this.destroyed) throw new Error('Compute backend is destroyed');
        if (!Number.isSafeInteger(size) || size < 0) {
            throw new Error(`Simulation buffer requires ${size} bytes; device storage/buffer limit is ${Number.MAX_SAFE_INTEGER}`);
        }
        return new Allocation(this, size);
    }

    uniforms() {
        if (this.destroyed) throw new Error('Compute backend is destroyed');
        return new Uniforms(this);
    }

    kernel(name) {
        const id = this.native.kernelId(name);
        const params = KERNEL_PARAMS[name];
        if (id < 0 || !params) throw new Error(`Unknown compute kernel ${name}`);
        return { name, id, params: new Set(params), destroy() {} };
    }

    pass(kernel, name) { return new Pass(this, kernel, name); }

    dispatch(sequence, label = 'softbody-step') {
        if (this.destroyed) throw new Error('Compute backend is destroyed');
        const launches = sequence.filter(pass => pass.groups[0] > 0);
        const bytes = new ArrayBuffer(Math.max(1, launches.length) * LAUNCH_BYTES);
        const view = new DataView(bytes);
        launches.forEach((pass, i) => pass.encode(view, i * LAUNCH_BYTES));
        for (const u of this.uniformSets) u.upload();
        if (launches.length) this.native.dispatch(this.handle, bytes, launches.length);
        this.lastLabel = label;
    }

    synchronize() { this.native.synchronize(this.handle); }

    destroy() {
        if (this.destroyed) return;
        for (const allocation of [...this.allocations]) allocation.destroy();
        this.uniformSets.clear();
        this.native.destroy(this.handle);
        this.destroyed = true;
    }
}
