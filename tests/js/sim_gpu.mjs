// Runs one scene through the JS Simulator on the HIP backend and writes what
// the pytest side compares with the oracle: node positions/velocities after
// the frames, the lattice the JS host built, and the deformed splats.
// Usage: node tests/js/sim_gpu.mjs <spec.json> <out-dir>
import fs from 'fs';
import path from 'path';
import { Simulator, SoftBody } from '../../visionary-sim_amd/js/simulator.mjs';
import { SplatDeformer } from '../../visionary-sim_amd/js/deformation/splats.mjs';
import { surfaceNodes } from '../../visionary-sim_amd/js/discretize/lattice.mjs';

function f32(file) {
    const b = fs.readFileSync(file);
    return new Float32Array(b.buffer, b.byteOffset, b.byteLength / 4);
}

function write(dir, name, typed) {
    fs.writeFileSync(path.join(dir, name), Buffer.from(typed.buffer, typed.byteOffset, typed.byteLength));
}

async function main() {
    const [specPath, out] = process.argv.slice(2);
    const spec = JSON.parse(fs.readFileSync(specPath, 'utf8'));
    fs.mkdirSync(out, { recursive: true });
    const body = new SoftBody({
        name: 'block',
        source: { kind: 'block', cells: spec.cells, origin: spec.origin, h: spec.h },
        material: spec.material,
        staticRegion: spec.pinned ? (x, y) => y <= spec.origin[1] + spec.h / 2 : null
    });
    const sim = new Simulator(0, [body], spec.tank);
    sim.clock.setStepDuration(spec.dt);
    let splats = null;
    if (spec.splats) {
        splats = new SplatDeformer(sim, 0, {
            means: f32(path.join(spec.splats, 'means.bin')),
            quats: f32(path.join(spec.splats, 'quats.bin')),
            scales: f32(path.join(spec.splats, 'scales.bin'))
        }, { axes: true });
    }
    if (spec.drive) {
        sim.setDriveLaw(spec.drive.law === 0 ? 'hard' : 'compliant');
        sim.hold(spec.drive.nodes, spec.drive.offsets, spec.drive.targets[0]);
    }
    const t0 = process.hrtime.bigint();
    for (let f = 0; f < spec.frames; f++) {
        if (spec.drive) sim.moveGrab(spec.drive.targets[f + 1]);
        sim.dispatch(spec.substeps);
        if (splats) splats.deform();
    }
    const pos = await sim.readPositions();
    const ms = Number(process.hrtime.bigint() - t0) / 1e6;
    const N = sim.particles.nodeCount;
    const planes = new Float32Array(3 * N);
    for (let n = 0; n < N; n++) for (let a = 0; a < 3; a++) planes[a * N + n] = pos[4 * n + a];
    write(out, 'pos.bin', planes);
    write(out, 'vel.bin', await sim.particles.readVelocityPlanes());
    const lat = sim.simBodies[0].proxy.lattice;
    write(out, 'rest.bin', lat.restPositions);
    write(out, 'w.bin', sim.particles.gpuInvMass);
    write(out, 'cells.bin', lat.cells);
    write(out, 'surface.bin', surfaceNodes(lat));
    if (splats) {
        const { means, cov, rotation, scale } = await splats.read();
        write(out, 'splat_means.bin', means);
        write(out, 'splat_cov.bin', cov);
        write(out, 'splat_rot.bin', rotation);
        write(out, 'splat_axis.bin', scale);
    }
    let grab = null;
    if (spec.grab) {
        // Pick and grip on the device against the positions written above.
        const hit = await sim.grab(spec.grab.origin, spec.grab.direction);
        const f32 = new Float32Array(sim.drive.words);
        const offsets = sim.drive.heldNodes.map((n, i) => [1, 2, 3].map(a => f32[a * sim.drive.capacity + i]));
        grab = { hit, held: sim.drive.heldNodes, offsets };
    }
    fs.writeFileSync(path.join(out, 'meta.json'), JSON.stringify({ nodes: N, ms, stats: sim.stats(), grab }));
    if (splats) splats.destroy();
    sim.destroy();
    console.log(`js host: ${N} nodes, ${spec.frames} frames in ${ms.toFixed(1)} ms`);
}

main().catch((error) => {
    console.error(error);
    process.exit(1);
});
