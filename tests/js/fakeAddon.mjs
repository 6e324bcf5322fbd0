// An in-memory stand-in for the Node-API addon (the role of the reference's
// test/helpers/fakeCompute.js): allocations are host byte arrays, launches
// are recorded, nothing executes.
const IDS = ['sim-predict', 'sim-finalize', 'sim-elastic-cells', 'sim-contact-bin', 'sim-contact-sort',
    'sim-contact-collide', 'sim-contact-settle', 'sim-drive-hard', 'sim-drive-compliant', 'sim-splat-deform',
    'sim-pick-ray', 'sim-grip-radius'];

export function fakeAddon() {
    let next = 256n;
    const memory = new Map();
    const api = {
        memory,
        launches: [],
        writes: 0,
        create(device) { if (device !== 0) throw new Error(`no HIP device ${device}`); return { device }; },
        destroy() { api.destroyed = true; },
        alloc(_h, bytes) {
            const handle = next;
            next += BigInt(Math.ceil(bytes / 256) * 256 + 256);
            memory.set(handle, new Uint8Array(bytes));
            return handle;
        },
        free(_h, handle) { memory.delete(handle); },
        write(_h, handle, offset, data) {
            const bytes = new Uint8Array(data.buffer, data.byteOffset, data.byteLength);
            const target = memory.get(handle);
            if (!target || Number(offset) + bytes.length > target.length) throw new Error('range exceeds buffer');
            target.set(bytes, Number(offset));
            api.writes++;
        },
        read(_h, handle, offset, view) {
            const source = memory.get(handle);
            view.set(source.subarray(Number(offset), Number(offset) + view.length));
        },
        kernelId(name) { return IDS.indexOf(name); },
        dispatch(_h, bytes, count) {
            const view = new DataView(bytes);
            for (let i = 0; i < count; i++) {
                const at = i * 112;
                const buffers = [];
                for (let s = 0; s < 10; s++) buffers.push(view.getBigUint64(at + 32 + 8 * s, true));
                api.launches.push({ kernel: IDS[view.getInt32(at, true)], grid: view.getUint32(at + 4, true),
                    block: view.getUint32(at + 8, true), plane: view.getUint32(at + 12, true),
                    flags: view.getUint32(at + 20, true), buffers });
            }
        },
        synchronize() {},
        launchSize() { return 112; },
        // Voxelisers: the call is recorded and a 2 x 1 x 1 block comes back.
        voxelCalls: [],
        voxelizeMesh(...args) { return api.voxelGrid('mesh', args); },
        voxelizeShape(...args) { return api.voxelGrid('shape', args); },
        voxelizeCloud(...args) { return api.voxelGrid('cloud', args); },
        voxelGrid(what, args) {
            api.voxelCalls.push({ what, args });
            return { dims: [2, 1, 1], origin: [0, 0, 0], h: 0.01, occupancy: Uint8Array.from([1, 1]) };
        }
    };
    return api;
}
