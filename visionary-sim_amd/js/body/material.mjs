// Material space (mirrors packages/visionary-sim/src/body/material.js:7-140):
// engineering constants an author sets, and the two conversions the solver needs.

/** nu is clamped to the range the stable Neo-Hookean solve holds (material.js:66). */
export const POISSON_RANGE = Object.freeze({ min: 0.25, max: 0.49 });

/** DEFAULT_MATERIAL (material.js:69-77). */
export const DEFAULT_MATERIAL = Object.freeze({
    youngsModulus: 1.2e5,
    poissonRatio: 0.40,
    density: 1000,
    friction: 0.4,
    damping: 0.6,
    relaxation: 0.012,
    maxStretch: 0
});

/** Zero disables the tensile limit; otherwise a finite ratio above one. */
export function validateMaxStretch(value = 0) {
    if (!Number.isFinite(value) || (value !== 0 && value <= 1)) {
        throw new RangeError('maxStretch must be zero (disabled) or a finite ratio greater than one');
    }
    return value;
}

/** Lame parameters with nu clamped and E floored at 1 (material.js:101-108). */
export function lameParameters({ youngsModulus, poissonRatio }) {
    const nu = Math.min(Math.max(poissonRatio, POISSON_RANGE.min), POISSON_RANGE.max);
    const e = Math.max(youngsModulus, 1);
    return { mu: e / (2 * (1 + nu)), lambda: (e * nu) / ((1 + nu) * (1 - 2 * nu)) };
}

/** Longitudinal wave speed sqrt((lambda + 2 mu) / rho) (material.js:137-140). */
export function waveSpeed(material) {
    const { mu, lambda } = lameParameters(material);
    return Math.sqrt((lambda + 2 * mu) / material.density);
}
