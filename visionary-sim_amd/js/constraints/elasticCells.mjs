// The elastic constraint over every lattice of the scene: one pass per
// non-empty colour, cells grouped by colour across bodies (the role of
// packages/visionary-sim/src/sim/discretize/lattice/elasticCells.js:126-268).
// Each colour owns a cell list (linear ids) and a tile buffer whose tiles stay
// inside one body, so a workgroup reads its body's constants once.
import { BODY, BODY_STRIDE } from '../native.mjs';
import { elementCompliance } from './compliance.mjs';
import { validateMaxStretch } from '../body/material.mjs';
import { COLOR_COUNT } from '../discretize/lattice.mjs';
import { tileWords } from '../state/particles.mjs';

/** Cells per colour-pass tile (one workgroup). */
export const CELL_TILE = 256;
const CELL_BLOCK = 256;

export function bindAccepted(pass, bindings) {
    for (const [name, value] of Object.entries(bindings)) {
        if (pass.kernel.params.has(name)) pass.set(name, value);
    }
}

/** Rest volume of one Kuhn tetrahedron (kuhn.js:16-18). */
export function tetRestVolume(h) {
    return (1 / 6) * Math.pow(h, 3);
}

export function elasticConstraints(backend, particles, proxies, materials) {
    const bodies = proxies.map((proxy, slot) => ({ lattice: proxy.lattice, material: materials[slot], slot }));
    return bodies.length > 0 ? [new ElasticCells(backend, particles, bodies)] : [];
}

export class ElasticCells {
    constructor(backend, particles, bodies) {
        this.particles = particles;
        this.bodies = bodies;
        this.shader = backend.kernel('sim-elastic-cells');
        const total = bodies.reduce((sum, { lattice }) => sum + lattice.cells.length, 0);
        const cells = new Uint32Array(Math.max(1, total));
        this.cellCount = total;
        this.colorRanges = [];
        const tilesPerColor = [];
        let cursor = 0;
        for (let color = 0; color < COLOR_COUNT; color++) {
            const start = cursor;
            const ranges = [];
            bodies.forEach(({ lattice, slot }) => {
                const range = lattice.colorRanges[color];
                cells.set(lattice.cells.subarray(range.start, range.start + range.count), cursor);
                ranges.push({ begin: cursor, count: range.count, body: slot });
                cursor += range.count;
            });
            this.colorRanges.push({ start, count: cursor - start });
            tilesPerColor.push(tileWords(ranges, CELL_TILE));
        }
        this.cells = backend.buffer(cells.byteLength);
        this.cells.write(0, cells);
        this.tiles = tilesPerColor.map(({ words }) => {
            const buffer = backend.buffer(words.byteLength);
            buffer.write(0, words);
            return buffer;
        });
        this.uploadParams();
        const batches = limitOnly => this.colorRanges
            .map((range, color) => ({ range, color }))
            .filter(({ range }) => range.count > 0)
            .map(({ color }) => {
                const pass = backend.pass(this.shader, `sim-elastic-${color}`);
                bindAccepted(pass, particles.bindings);
                pass.set('cells', this.cells);
                pass.set('tiles', this.tiles[color]);
                pass.set('limitOnly', limitOnly);
                pass.set('block', CELL_BLOCK);
                pass.setupDispatch(tilesPerColor[color].count);
                return pass;
            });
        this.batches = batches(0);
        this.stretchBatches = batches(1);
        this.limitedPasses = [...this.batches, ...this.stretchBatches, ...this.stretchBatches, ...this.stretchBatches];
    }

    get passes() {
        return this.bodies.some(({ material }) => material.maxStretch > 1) ? this.limitedPasses : this.batches;
    }

    /** Rewrites the elastic half of every body record (elasticCells.js:222-249). */
    uploadParams() {
        const { bodyF32, bodyU32 } = this.particles;
        for (const { lattice, material, slot } of this.bodies) {
            const { cellDims, nodeDims, h, origin } = lattice;
            const { deviatoric, hydrostatic, gamma } = elementCompliance(material, tetRestVolume(h));
            const base = slot * (BODY_STRIDE / 4);
            for (let a = 0; a < 3; a++) {
                bodyU32[base + BODY.cellDims + a] = cellDims[a];
                bodyU32[base + BODY.nodeDims + a] = nodeDims[a];
                bodyF32[base + BODY.origin + a] = origin[a];
            }
            bodyF32[base + BODY.h] = h;
            bodyF32[base + BODY.alphaDev] = deviatoric;
            bodyF32[base + BODY.alphaHyd] = hydrostatic;
            bodyF32[base + BODY.gamma] = gamma;
            bodyF32[base + BODY.relaxation] = material.relaxation;
            bodyF32[base + BODY.maxStretch] = validateMaxStretch(material.maxStretch);
        }
        this.particles.uploadParams();
    }

    /** dt lives in the shared uniforms; nothing per pass. */
    setSubstepDt() {}

    destroy() {
        this.cells.destroy();
        for (const t of this.tiles) t.destroy();
        for (const pass of [...this.batches, ...this.stretchBatches]) pass.destroy();
    }
}
