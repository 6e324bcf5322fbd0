// The simulation's public surface, kept from the kernel package
// (packages/visionary-sim/src/sim/simulator.js:35-518): assembly in
// dependency order, step(dt) through the fixed clock, dispatch(steps) as one
// submission per frame, stats, material changes, reset, readback. The
// compute backend underneath is the HIP one. grab(origin, direction) picks
// and grips on the device (grab.mjs); hold() takes a held set directly.
import { ComputeBackend } from './compute/backend.mjs';
import { proxyFor } from './discretize/lattice.mjs';
import { Particles } from './state/particles.mjs';
import { KinematicDrive } from './constraints/kinematicDrive.mjs';
import { elasticConstraints } from './constraints/elasticCells.mjs';
import { BodyContact } from './constraints/bodyContact.mjs';
import { Stepper } from './solver/stepper.mjs';
import { FixedTimestep, stableStepDuration } from './solver/timestep.mjs';
import { collapseScale, gammaOf, waveRatio } from './constraints/compliance.mjs';
import { DEFAULT_MATERIAL, validateMaxStretch } from './body/material.mjs';
import { DEFAULT_GRIPPED_PART_RADIUS, GRIP_STRATEGIES, GrabQueries, bodyOfNode } from './grab.mjs';

const DEFAULT_PARAMS = { gravity: [0, -9.81, 0], paused: false };

/**
 * A body as an author describes it (body/softBody.js:37-60). `source` is
 * rest geometry: { kind: 'mesh', mesh }, { kind: 'shape', shape },
 * { kind: 'cloud', cloud } (voxelised on the device at `elementSize`), or an
 * occupancy grid ({ kind: 'lattice', grid }) or block ({ kind: 'block', ... }).
 */
export class SoftBody {
    constructor({ name, source, material = {}, staticRegion = null, velocity = null, elementSize = 0.01 }) {
        this.name = name;
        this.source = source;
        this.material = Object.assign({}, DEFAULT_MATERIAL, material);
        validateMaxStretch(this.material.maxStretch);
        this.staticRegion = staticRegion;
        this.velocity = velocity;
        this.elementSize = elementSize;
    }
}

/** Body-local node ids whose rest position the region claims (staticPrescription.js:9-20). */
export function staticNodesInRegion(restPositions, region = null) {
    const nodes = [];
    if (region === null) return nodes;
    for (let n = 0; n < restPositions.length / 3; n++) {
        if (region(restPositions[n * 3], restPositions[n * 3 + 1], restPositions[n * 3 + 2])) nodes.push(n);
    }
    return nodes;
}

export class Simulator {
    /**
     * @param {number|ComputeBackend} device - HIP device index, or a backend to borrow.
     * @param {SoftBody[]} bodies - Bodies to simulate.
     * @param {{ min: number[], max: number[] }} tank - Tank interior bounds.
     * @param {{ bucketSize?: number }} [options] - Contact grid override.
     */
    constructor(device, bodies, tank, options = {}) {
        const owned = !(device instanceof ComputeBackend);
        const backend = owned ? new ComputeBackend(device) : device;
        this.backend = backend;
        this.ownsBackend = owned;
        this.destroyed = false;
        try {
            this.tank = tank;
            this.params = Object.assign({}, DEFAULT_PARAMS);
            this.simBodies = bodies.map(body => ({ body, proxy: proxyFor(body, backend) }));
            this.uniforms = backend.uniforms();
            this.particles = new Particles(backend, this.simBodies.map(({ body, proxy }) => Object.assign({}, proxy.dof, {
                staticNodes: staticNodesInRegion(proxy.dof.restPositions, body.staticRegion),
                velocity: body.velocity
            })), this.uniforms);
            this.drive = new KinematicDrive(backend, this.particles);
            this.elasticSets = elasticConstraints(backend, this.particles,
                this.simBodies.map(({ proxy }) => proxy), this.simBodies.map(({ body }) => body.material));
            this.contact = new BodyContact(backend, this.particles, this.simBodies.map(({ proxy }) => proxy), tank, options);
            // Solve order (simulator.js:76-83): drive, elasticity, contact.
            this.stepper = new Stepper(backend, this.particles, [this.drive, ...this.elasticSets, this.contact]);
            this.clock = new FixedTimestep(stableStepDuration(this.waveRate()));
            this.simBodies.forEach(({ body }, b) => {
                this.particles.setParams(b, { damping: body.material.damping, friction: body.material.friction });
            });
            this.grabTarget = { position: [0, 0, 0], previousPosition: [0, 0, 0] };
            this.grabQueries = new GrabQueries(backend, this.particles, this.drive.capacity);
            this.grip = { strategy: 'withinRadius', radius: DEFAULT_GRIPPED_PART_RADIUS };
            this.grabGeneration = 0;
            this.pendingStep = false;
        } catch (error) {
            if (owned) backend.destroy();
            throw error;
        }
    }

    worst() {
        if (this.simBodies.length === 0) return null;
        return this.simBodies.reduce((worst, entry) =>
            (waveRatio(entry.body.material, entry.proxy.cflLength) > waveRatio(worst.body.material, worst.proxy.cflLength)
                ? entry : worst));
    }

    waveRate() {
        const limiting = this.worst();
        return limiting ? waveRatio(limiting.body.material, limiting.proxy.cflLength) : 0;
    }

    cflNumber() { return this.waveRate() * this.clock.stepDuration; }

    stats() {
        const timing = this.clock.snapshot();
        const worst = this.worst();
        return Object.assign({
            bodyCount: this.simBodies.length,
            nodes: this.particles.nodeCount,
            cells: this.elasticSets.reduce((sum, set) => sum + set.cellCount, 0),
            contactNodes: this.contact.surfaceCount,
            contact: this.contact.enabled,
            cfl: this.cflNumber(),
            limitedBy: worst ? worst.body.name : 'nothing',
            perBody: this.simBodies.map(({ body }) => {
                const gamma = gammaOf(body.material);
                return { name: body.name, gamma, collapseScale: collapseScale(gamma) };
            })
        }, timing);
    }

    /**
     * Advances by one render frame of `dt` seconds (simulator.js:239-278).
     * A scene with no nodes has nothing to advance: the clock idles. Input
     * that is not a finite non-negative number counts as a frame of no time.
     * A step asked for with requestStep() is taken in this frame: paused, as
     * the one step of the frame; running, as a step forced on top of the
     * frame's banked time. Paused with no request, the clock idles.
     */
    step(dt) {
        const requested = this.pendingStep;
        this.pendingStep = false;
        if (this.particles.nodeCount === 0) {
            this.clock.idle();
            return;
        }
        const frameTime = Number.isFinite(dt) && dt >= 0 ? dt : 0;
        let timing = null;
        if (!this.params.paused) timing = this.clock.consume(frameTime, { forceStep: requested });
        else if (requested) timing = this.clock.singleStep();
        if (timing === null) {
            this.clock.idle();
            return;
        }
        this.dispatch(timing.steps);
    }

    requestStep() { this.pendingStep = true; }

    /**
     * All `steps` substeps of a frame in one submission (simulator.js:292-318).
     * The uniforms of the frame are its substep, gravity and tank. When
     * something is held, the drive is told where the grab target is and how
     * fast it moved: its displacement since the last dispatch spread over
     * this frame's simulated time.
     */
    dispatch(steps) {
        if (!(steps >= 1) || this.particles.nodeCount === 0) return;
        const substepDt = this.clock.stepDuration;
        this.stepper.setFrameParameters({ substepDt, gravity: this.params.gravity, tankMin: this.tank.min,
            tankMax: this.tank.max });
        if (this.drive.drivenCount > 0) this.aimDrive(steps * substepDt);
        this.backend.dispatch(this.stepper.sequence(steps), 'softbody-step');
    }

    /** The drive's target for a frame of `frameDt` seconds; the target's last position moves on. */
    aimDrive(frameDt) {
        const target = this.grabTarget;
        const velocity = target.position.map((x, axis) => (x - target.previousPosition[axis]) / frameDt);
        this.drive.setGrabTarget(target.position, velocity);
        target.previousPosition = target.position;
    }

    /**
     * What a line through the scene hits, holding it if it hit something
     * (simulator.js:366-418). The pick and the grip run on the device; only
     * the anchor, its position and the held set come back.
     *
     * @param {number[]} origin - Ray origin in world space.
     * @param {number[]} direction - Unit ray direction.
     * @returns {Promise<{ bodyIndex: number, position: number[], held: number }|null>}
     */
    async grab(origin, direction) {
        if (this.destroyed) return null;
        const generation = ++this.grabGeneration;
        // Aim tolerance: a little over the widest node spacing.
        const aimRadius = this.coarsestSpacing() * 1.2;
        const { node: anchor } = await this.grabQueries.pick(origin, direction, aimRadius);
        if (generation !== this.grabGeneration || this.destroyed) return null;
        if (anchor < 0) return null;
        let grip;
        if (this.grip.strategy === 'singleNode') {
            // The pick only returns grabbable nodes, so the anchor is held.
            grip = { nodes: [anchor], offsets: [[0, 0, 0]], found: 1 };
        } else {
            grip = await this.grabQueries.grip(anchor, {
                radius: this.grip.radius, nodeOffsets: this.particles.offsets, limit: this.drive.capacity
            });
        }
        if (generation !== this.grabGeneration || this.destroyed) return null;
        const held = grip.nodes.length;
        if (held < grip.found) {
            console.warn(`[grip] ${grip.found} nodes within ${this.grip.radius} m, only the nearest ` +
                `${held} held -- this grab's extent is resolution-dependent again.`);
        }
        // The anchor's live position is where the hand starts: the target
        // begins there, at rest.
        const position = await this.grabQueries.position(anchor);
        this.hold(grip.nodes, grip.offsets, position);
        return { bodyIndex: bodyOfNode(this.particles.offsets, anchor), position, held };
    }

    /** Widest node spacing of any body, the aim tolerance's scale (simulator.js:420-423). */
    coarsestSpacing() {
        return this.simBodies.reduce((coarsest, { proxy }) => Math.max(coarsest, proxy.spacing.max), 0);
    }

    /** Closest node spacing of any body (simulator.js:425-435). */
    finestSpacing() {
        return this.simBodies.reduce((finest, { proxy }) => Math.min(finest, proxy.spacing.min), Infinity);
    }

    /** `'singleNode'` or `'withinRadius'` (simulator.js:437-447). */
    setGrip(strategy) {
        if (!GRIP_STRATEGIES.includes(strategy)) throw new Error(`[grip] unknown strategy "${strategy}"`);
        this.grip.strategy = strategy;
    }

    /** World-space hand radius for `withinRadius`, metres. */
    setGripRadius(radius) { this.grip.radius = radius; }

    /** Holds `nodes` (global ids) at `offsets` from an anchor at `position`. */
    hold(nodes, offsets, position) {
        this.drive.hold(nodes, offsets);
        this.grabTarget.position = position.slice();
        this.grabTarget.previousPosition = position.slice();
    }

    moveGrab(position) { this.grabTarget.position = position.slice(); }

    /** Lets go of the held set; a grab still in flight is abandoned. */
    releaseGrab() {
        this.grabGeneration++;
        if (!this.destroyed) this.drive.release();
    }

    setDriveLaw(law) { this.drive.law = law; }

    /** The drive law in use (simulator.js:487). */
    get driveLaw() { return this.drive.law; }

    setDriveResponseTime(seconds) { this.drive.responseTime = seconds; }

    setContact(enabled) { this.contact.enabled = enabled; }

    /**
     * A material change (simulator.js:332-353). Density rescales the body's
     * node masses (they belong to the discretisation); damping and friction
     * are node uniforms; the rest are elastic uniforms. A change of
     * stiffness or density changes the wave rate, so the clock's step is
     * re-derived, its banked seconds kept.
     */
    setMaterial(bodyIndex, changes) {
        if ('maxStretch' in changes) validateMaxStretch(changes.maxStretch);
        const { body, proxy } = this.simBodies[bodyIndex];
        const material = body.material;
        for (const key of Object.keys(changes)) {
            if (key !== 'density') material[key] = changes[key];
        }
        const density = changes.density;
        if (density > 0 && density !== material.density) {
            proxy.setDensity(material.density, density);
            material.density = density;
            this.particles.uploadMasses();
        }
        this.particles.setParams(bodyIndex, { damping: material.damping, friction: material.friction });
        this.elasticSets.forEach(set => set.uploadParams());
        this.clock.setStepDuration(stableStepDuration(this.waveRate()));
    }

    /** A body's deformation field: the live node buffer and where the body sits in it. */
    deformationOf(bodyIndex) {
        const { proxy } = this.simBodies[bodyIndex];
        const { origin, h, cellDims, nodeDims } = proxy.lattice;
        return {
            kernel: 'sim-splat-deform',
            parameters: {
                deformNodePos: this.particles.nodePos, deformOrigin: origin, deformH: h,
                deformCellDims: cellDims, deformNodeDims: nodeDims, deformNodeOffset: this.particles.offsetOf(bodyIndex),
                body: bodyIndex
            }
        };
    }

    /** Every body back to rest, nothing held, the clock at zero (simulator.js:495-500). */
    reset() {
        this.pendingStep = false;
        this.releaseGrab();
        this.clock.reset();
        this.particles.reset();
    }

    /** Live node positions; each call returns its own array (simulator.js:502-503). */
    readPositions() { return this.particles.readPositions(); }

    destroy() {
        if (this.destroyed) return;
        this.destroyed = true;
        this.grabGeneration++;
        this.grabQueries.destroy();
        this.stepper.destroy();
        this.contact.destroy();
        for (const elastic of this.elasticSets) elastic.destroy();
        this.drive.destroy();
        this.particles.destroy();
        if (this.ownsBackend) this.backend.destroy();
    }
}
