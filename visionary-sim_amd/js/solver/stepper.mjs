// The substep loop (the role of packages/visionary-sim/src/sim/solver/
// stepper.js:198-275): predict, every constraint's passes in solve order,
// finalize; all substeps of a frame in one dispatch.
import { bindAccepted } from '../constraints/elasticCells.mjs';

const NODE_BLOCK = 256;

export class Stepper {
    constructor(backend, particles, constraints) {
        this.particles = particles;
        this.constraints = constraints;
        const make = name => {
            const pass = backend.pass(backend.kernel(name), name);
            bindAccepted(pass, particles.bindings);
            pass.set('tiles', particles.tiles);
            pass.set('block', NODE_BLOCK);
            pass.setupDispatch(particles.tileCount);
            return pass;
        };
        this.predict = make('sim-predict');
        this.finalize = make('sim-finalize');
        this.own = [this.predict, this.finalize];
    }

    setFrameParameters({ substepDt, gravity, tankMin, tankMax }) {
        const u = this.particles.uniforms;
        u.set('dt', substepDt);
        u.set('gravity', gravity);
        u.set('tankMin', tankMin);
        u.set('tankMax', tankMax);
        this.particles.setSubstepDt(substepDt);
        for (const constraint of this.constraints) constraint.setSubstepDt(substepDt);
    }

    sequence(steps) {
        const projections = this.constraints.flatMap(constraint => constraint.passes);
        const list = [];
        for (let s = 0; s < steps; s++) list.push(this.predict, ...projections, this.finalize);
        return list;
    }

    destroy() {
        for (const pass of this.own) pass.destroy();
    }
}
