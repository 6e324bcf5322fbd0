// A capture through the JS host on the HIP device: PLY bytes -> loadCloud ->
// SoftBody({ kind: 'cloud' }) voxelised on the device -> lattice -> frames.
// Writes the lattice and the final node state for the pytest side.
// Usage: node tests/js/capture_gpu.mjs <spec.json> <out-dir>
import fs from 'fs';
import path from 'path';
import { Simulator, SoftBody } from '../../visionary-sim_amd/js/simulator.mjs';
import { loadCloud, restCloud } from '../../visionary-sim_amd/js/assets/ply.mjs';

function write(dir, name, typed) {
    fs.writeFileSync(path.join(dir, name), Buffer.from(typed.buffer, typed.byteOffset, typed.byteLength));
}

async function main() {
    const [specPath, out] = process.argv.slice(2);
    const spec = JSON.parse(fs.readFileSync(specPath, 'utf8'));
    fs.mkdirSync(out, { recursive: true });
    const cloud = loadCloud(spec.ply);
    const body = new SoftBody({ name: 'capture', source: { kind: 'cloud', cloud: restCloud(cloud) },
                                material: spec.material, elementSize: spec.elementSize });
    const sim = new Simulator(0, [body], spec.tank);
    sim.clock.setStepDuration(spec.dt);
    for (let f = 0; f < spec.frames; f++) sim.dispatch(spec.substeps);
    const pos = await sim.readPositions();
    const N = sim.particles.nodeCount;
    const planes = new Float32Array(3 * N);
    for (let n = 0; n < N; n++) for (let a = 0; a < 3; a++) planes[a * N + n] = pos[4 * n + a];
    write(out, 'pos.bin', planes);
    const lat = sim.simBodies[0].proxy.lattice;
    write(out, 'rest.bin', lat.restPositions);
    write(out, 'cells.bin', lat.cells);
    write(out, 'w.bin', sim.particles.gpuInvMass);
    fs.writeFileSync(path.join(out, 'meta.json'), JSON.stringify({ nodes: N, cellDims: lat.cellDims,
                                                                   origin: lat.origin, h: lat.h }));
    sim.destroy();
    console.log(`capture: ${N} nodes, ${lat.cells.length} cells`);
}

main().catch((error) => {
    console.error(error);
    process.exit(1);
});
