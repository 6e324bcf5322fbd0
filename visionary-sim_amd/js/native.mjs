// The Node-API addon over libvsim.so (visionary-sim_amd/node/addon.cpp).
// Loaded once; a missing build is an error here, never a silent fallback.
import { createRequire } from 'module';
import { fileURLToPath } from 'url';
import path from 'path';
import fs from 'fs';

const here = path.dirname(fileURLToPath(import.meta.url));
export const ADDON_PATH = path.join(here, '..', 'lib', 'vsim.node');

let loaded = null;

/** @returns {object} The addon's exports. */
export function addon() {
    if (loaded) return loaded;
    if (!fs.existsSync(ADDON_PATH)) {
        throw new Error(`${ADDON_PATH} is not built; run python visionary-sim_amd/build.py`);
    }
    loaded = createRequire(import.meta.url)(ADDON_PATH);
    return loaded;
}

/** Buffer slots of a launch (VS_B_* in include/vsim.h). */
export const SLOT = Object.freeze({
    nodePos: 0, nodePrev: 1, nodeVel: 2, nodeInvMass: 3, bodies: 4, params: 5, tiles: 6,
    items: 7, contactGrid: 8, out: 9
});

/** Size of one vs_launch record. */
export const LAUNCH_BYTES = 112;

/** vs_body: 80 bytes, word offsets. */
export const BODY_STRIDE = 80;
export const BODY = Object.freeze({
    cellDims: 0, nodeOffset: 3, nodeDims: 4, scene: 7, h: 8, alphaDev: 9, alphaHyd: 10, gamma: 11,
    relaxation: 12, maxStretch: 13, friction: 14, damp: 15, origin: 16, nodeCount: 19
});

/** vs_scene_params: 160 bytes, word offsets and types. */
export const SCENE_BYTES = 160;
export const SCENE = Object.freeze({
    dt: [0, 'f32', 1], gravity: [1, 'f32', 3], tankMin: [4, 'f32', 3], gain: [7, 'f32', 1],
    tankMax: [8, 'f32', 3], bucketSize: [11, 'f32', 1], gridOrigin: [12, 'f32', 3], bucketCount: [15, 'u32', 1],
    gridDims: [16, 'u32', 3], gridBase: [19, 'u32', 1], grabTarget: [20, 'f32', 3], responseTime: [23, 'f32', 1],
    grabTargetVelocity: [24, 'f32', 3], driveLaw: [27, 'u32', 1], surfBase: [28, 'u32', 1],
    surfaceCount: [29, 'u32', 1], bucketSlots: [30, 'u32', 1], flags: [31, 'u32', 1], sphere: [32, 'f32', 4],
    reachMax: [36, 'f32', 1], skipMargin: [37, 'f32', 1]
});

/** vs_tile: begin, count, body, scene (16 bytes); a tile buffer starts with a 16-byte header. */
export const TILE_WORDS = 4;
