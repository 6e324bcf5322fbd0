// The PLY container (apps/simulator/src/assets/splats/ply.js:65-152) and the
// Gaussian cloud read out of it (apps/simulator/src/assets/splats/cloud.js:
// 28-81), parsed natively on the host (visionary-sim_amd/csrc/ply.cpp) with
// the reference's layouts, conversions and error messages.
import fs from 'fs';
import { addon } from '../native.mjs';

/** The properties a 3DGS PLY must carry (cloud.js:28-34). */
export const REQUIRED_PROPERTIES = [
    'x', 'y', 'z',
    'f_dc_0', 'f_dc_1', 'f_dc_2',
    'opacity',
    'scale_0', 'scale_1', 'scale_2',
    'rot_0', 'rot_1', 'rot_2', 'rot_3'
];

/**
 * readPly: the first element of a binary little-endian PLY.
 * @param {ArrayBuffer|Uint8Array} buffer - The whole file.
 * @param {{ properties?: string[], native?: object }} [options] - Names to keep (all by default).
 * @returns {{ count: number, properties: Record<string, Float32Array> }} The element.
 */
export function readPly(buffer, { properties, native } = {}) {
    const lib = native || addon();
    const names = properties || lib.plyProperties(buffer).names;
    return lib.readPly(buffer, names);
}

/**
 * cloudFromPly: positions, xyzw unit rotations (w >= 0), scales in metres,
 * opacity in [0, 1], colours from the degree-0 harmonic.
 * @param {{ count: number, properties: Record<string, Float32Array> }} ply - An element holding splats.
 * @param {{ native?: object }} [options] - Addon stand-in for tests.
 */
export function cloudFromPly({ count, properties }, { native } = {}) {
    const lib = native || addon();
    return lib.cloudFromColumns(count, REQUIRED_PROPERTIES.map(name => properties[name]));
}

/**
 * loadCloud (cloud.js:40-46) from a file path or the bytes of one (the
 * reference fetches a URL; this host reads files).
 */
export function loadCloud(source, options = {}) {
    const bytes = typeof source === 'string' ? fs.readFileSync(source) : source;
    return cloudFromPly(readPly(bytes, Object.assign({}, options, { properties: REQUIRED_PROPERTIES })), options);
}

/** restCloud (cloud.js:127-129): the physical Gaussian data only. */
export function restCloud({ count, positions, rotations, scales, opacity }) {
    return { count, positions, rotations, scales, opacity };
}
