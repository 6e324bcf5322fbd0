// Nodes that are told, not solved for (the role of packages/visionary-sim/
// src/sim/constraints/kinematicDrive.js:186-394): a hard law that prescribes
// the held nodes and moves them itself, and a compliant law that nudges them.
// Driven records are four planes of MAX_DRIVEN words: node, offset x, y, z.
import { tileWords } from '../state/particles.mjs';
import { bindAccepted } from './elasticCells.mjs';

export const MAX_DRIVEN = 4096;
const RESPONSE_TIME = 0.05;
const DRIVE_TILE = 64;

export class KinematicDrive {
    constructor(backend, particles) {
        this.backend = backend;
        this.particles = particles;
        this._law = 'hard';
        this.heldNodes = [];
        this.drivenCount = 0;
        this.driven = backend.buffer(4 * 4 * MAX_DRIVEN);
        this.words = new ArrayBuffer(4 * 4 * MAX_DRIVEN);
        this.tiles = backend.buffer(16 * (1 + Math.ceil(MAX_DRIVEN / DRIVE_TILE)));
        this.computes = {};
        for (const law of ['hard', 'compliant']) {
            const pass = backend.pass(backend.kernel(`sim-drive-${law}`), `sim-drive-${law}`);
            bindAccepted(pass, particles.bindings);
            pass.set('driven', this.driven);
            pass.set('items', MAX_DRIVEN);
            pass.set('tiles', this.tiles);
            pass.set('block', DRIVE_TILE);
            pass.setupDispatch(0);
            this.computes[law] = pass;
        }
        this.responseTime = RESPONSE_TIME;
        this.particles.uniforms.set('driveLaw', 0);
    }

    set responseTime(seconds) {
        if (!Number.isFinite(seconds) || seconds <= 0) {
            throw new RangeError('drive responseTime must be a finite positive number of seconds');
        }
        this._responseTime = seconds;
        this.particles.uniforms.set('responseTime', seconds);
    }

    get responseTime() { return this._responseTime; }

    get capacity() { return MAX_DRIVEN; }

    hold(nodes, offsets) {
        if (nodes.some(node => this.particles.staticPrescribed.has(node))) {
            throw new Error('[drive] cannot hold a static prescribed node');
        }
        if (nodes.length === 0) {
            this.release();
            return;
        }
        const count = Math.min(nodes.length, MAX_DRIVEN);
        const u32 = new Uint32Array(this.words);
        const f32 = new Float32Array(this.words);
        for (let i = 0; i < count; i++) {
            u32[i] = nodes[i];
            for (let a = 0; a < 3; a++) f32[(1 + a) * MAX_DRIVEN + i] = offsets[i][a];
        }
        this.driven.write(0, new Uint8Array(this.words));
        const { words, count: tiles } = tileWords([{ begin: 0, count, body: 0 }], DRIVE_TILE);
        this.tiles.write(0, words);
        for (const pass of Object.values(this.computes)) pass.setupDispatch(tiles);
        this.drivenCount = count;
        this.heldNodes = nodes.slice(0, count);
        this.applyPrescription();
    }

    setGrabTarget(target, velocity) {
        this.particles.uniforms.set('grabTarget', target);
        this.particles.uniforms.set('grabTargetVelocity', velocity);
    }

    release() {
        this.drivenCount = 0;
        this.heldNodes = [];
        for (const pass of Object.values(this.computes)) pass.setupDispatch(0);
        this.particles.clearPrescribed();
    }

    set law(name) {
        if (name !== 'hard' && name !== 'compliant') throw new Error(`[drive] unknown law "${name}"`);
        this._law = name;
        this.particles.uniforms.set('driveLaw', name === 'hard' ? 0 : 1);
        this.applyPrescription();
    }

    get law() { return this._law; }

    applyPrescription() {
        if (this.drivenCount > 0 && this._law === 'hard') this.particles.setPrescribed(this.heldNodes);
        else this.particles.clearPrescribed();
    }

    get passes() { return this.drivenCount > 0 ? [this.computes[this._law]] : []; }

    setSubstepDt() {}

    destroy() {
        this.driven.destroy();
        this.tiles.destroy();
    }
}
