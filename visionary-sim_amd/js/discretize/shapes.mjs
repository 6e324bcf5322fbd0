// The analytic shapes the device voxeliser evaluates, with the reference's
// signed distances (apps/simulator/src/assets/shapes.js:32-156): each carries
// its bounds, its sdf as a JS function, and a `device` record (kind, size,
// translations) that voxelizeSdf sends to the GPU. Render meshes
// (`surface`) are not part of the simulation and are not built here.

/** Most translations a device shape carries (vs_shape.offset). */
export const MAX_TRANSLATIONS = 4;

/** torus (shapes.js:32-58): ring in the xz plane. */
export function torus({ majorRadius, minorRadius }) {
    const outer = majorRadius + minorRadius;
    return {
        bounds: { min: [-outer, -minorRadius, -outer], max: [outer, minorRadius, outer] },
        sdf: (x, y, z) => Math.hypot(Math.hypot(x, z) - majorRadius, y) - minorRadius,
        device: { kind: 'torus', size: [majorRadius, minorRadius, 0], offsets: [] }
    };
}

/** sphere (shapes.js:60-81). */
export function sphere({ radius }) {
    return {
        bounds: { min: [-radius, -radius, -radius], max: [radius, radius, radius] },
        sdf: (x, y, z) => Math.hypot(x, y, z) - radius,
        device: { kind: 'sphere', size: [radius, 0, 0], offsets: [] }
    };
}

/** box (shapes.js:89-110). */
export function box({ halfExtents }) {
    const [hx, hy, hz] = halfExtents;
    return {
        bounds: { min: [-hx, -hy, -hz], max: [hx, hy, hz] },
        sdf: (x, y, z) => {
            const dx = Math.abs(x) - hx;
            const dy = Math.abs(y) - hy;
            const dz = Math.abs(z) - hz;
            const outside = Math.hypot(Math.max(dx, 0), Math.max(dy, 0), Math.max(dz, 0));
            return outside + Math.min(Math.max(dx, dy, dz), 0);
        },
        device: { kind: 'box', size: [hx, hy, hz], offsets: [] }
    };
}

/**
 * translateShape (shapes.js:148-160). The device record keeps the chain of
 * translations (at most four); a longer chain can no longer run on the device.
 */
export function translateShape(shape, offset) {
    const [dx, dy, dz] = offset;
    const d = shape.device;
    return {
        bounds: {
            min: [shape.bounds.min[0] + dx, shape.bounds.min[1] + dy, shape.bounds.min[2] + dz],
            max: [shape.bounds.max[0] + dx, shape.bounds.max[1] + dy, shape.bounds.max[2] + dz]
        },
        sdf: (x, y, z) => shape.sdf(x - dx, y - dy, z - dz),
        device: d && d.offsets.length < MAX_TRANSLATIONS
            ? { kind: d.kind, size: d.size.slice(), offsets: d.offsets.concat([[dx, dy, dz]]) }
            : null
    };
}
