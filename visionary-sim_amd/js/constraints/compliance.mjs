// XPBD constants of the stable Neo-Hookean pair (mirrors
// packages/visionary-sim/src/sim/constraints/compliance.js:30-109).
import { lameParameters, waveSpeed } from '../body/material.mjs';

/** 1/(mu V), 1/(lambda V) and gamma = 1 + mu/lambda, in f64 on the host. */
export function elementCompliance(material, restVolume) {
    const { mu, lambda } = lameParameters(material);
    return { deviatoric: 1 / (mu * restVolume), hydrostatic: 1 / (lambda * restVolume), gamma: gammaOf(material) };
}

export function gammaOf(material) {
    const { mu, lambda } = lameParameters(material);
    return 1 + mu / lambda;
}

/** The uniform-scale barrier: root of s^3 + s^2 + s = gamma - 1 by bisection. */
export function collapseScale(gamma) {
    let low = 0;
    let high = 4;
    for (let i = 0; i < 40; i++) {
        const s = (low + high) / 2;
        if (s * s * s + s * s + s < gamma - 1) low = s; else high = s;
    }
    return (low + high) / 2;
}

/** Wave crossings of `cflLength` per second. */
export function waveRatio(material, cflLength) {
    return waveSpeed(material) / cflLength;
}
