// The fixed physical timestep of the kernel package's solver interface
// (packages/visionary-sim/src/sim/solver/timestep.js:17-180): the exported
// names, constants, error messages and snapshot fields are that interface.
// The behaviour, as specified:
//
//  * Frame time handed to consume() goes into a bank. It leaves the bank only
//    as whole steps of the fixed duration; a tolerance of 1e-9 of a step
//    lets summed frame times that fall a rounding error short still make the
//    step (1/60 s banked at 1/600 s steps is 10 steps, not 9).
//  * A call runs at most maxStepsPerFrame steps. Whole steps beyond that
//    leave the bank unrun and are counted as dropped time, so after any call
//    the bank holds less than one step: one slow frame never makes the next
//    one slower, and a step is never enlarged to catch up.
//  * forceStep runs one step when none was due, without taking it from the
//    bank. singleStep() runs one step with no frame time; idle() is a frame
//    that runs nothing and banks nothing.
//  * Changing the step duration keeps the banked seconds.
//  * snapshot() reports the last frame (steps, elapsed, dropped) and the
//    running totals (steps, simulated and dropped seconds, the bank).

export const TARGET_CFL = 0.75;
export const MAX_STEP_DURATION = 1 / 240;
export const MAX_STEPS_PER_FRAME = 24;

const STEP_TOLERANCE = 1e-9;

export function stableStepDuration(waveRate) {
    if (!(Number.isFinite(waveRate) && waveRate >= 0)) {
        throw new RangeError(`waveRate must be a finite non-negative number, got ${waveRate}`);
    }
    if (waveRate === 0) return MAX_STEP_DURATION;
    return Math.min(MAX_STEP_DURATION, TARGET_CFL / waveRate);
}

function requireStepDuration(value) {
    if (!(Number.isFinite(value) && value > 0)) {
        throw new RangeError(`stepDuration must be a finite positive number, got ${value}`);
    }
    return value;
}

/**
 * How a bank of `seconds` is spent in steps of `step` with at most `budget`
 * of them run: { run, dropped, remainder }, remainder the seconds kept.
 */
function planSteps(seconds, step, budget) {
    const tolerance = step * STEP_TOLERANCE;
    const whole = Math.floor((seconds + tolerance) / step);
    let remainder = seconds - whole * step;
    if (remainder < 0 && remainder > -tolerance) remainder = 0;
    const run = Math.min(whole, budget);
    return { run, dropped: whole - run, remainder };
}

export class FixedTimestep {
    constructor(stepDuration, options = {}) {
        const { maxStepsPerFrame = MAX_STEPS_PER_FRAME } = options;
        this.stepDuration = requireStepDuration(stepDuration);
        if (!(Number.isInteger(maxStepsPerFrame) && maxStepsPerFrame >= 1)) {
            throw new RangeError(`maxStepsPerFrame must be a positive integer, got ${maxStepsPerFrame}`);
        }
        this.maxStepsPerFrame = maxStepsPerFrame;
        this.reset();
    }

    setStepDuration(stepDuration) {
        this.stepDuration = requireStepDuration(stepDuration);
    }

    consume(elapsed, options = {}) {
        if (!(Number.isFinite(elapsed) && elapsed >= 0)) {
            throw new RangeError(`elapsed must be a finite non-negative number, got ${elapsed}`);
        }
        const plan = planSteps(this.accumulator + elapsed, this.stepDuration, this.maxStepsPerFrame);
        this.accumulator = plan.remainder;
        const steps = plan.run === 0 && options.forceStep ? 1 : plan.run;
        return this.record(steps, elapsed, plan.dropped * this.stepDuration);
    }

    idle() {
        return this.record(0, 0, 0);
    }

    singleStep() {
        return this.record(1, 0, 0);
    }

    reset() {
        this.accumulator = 0;
        this.simulatedTime = 0;
        this.totalSteps = 0;
        this.droppedTime = 0;
        this.lastSteps = 0;
        this.lastElapsed = 0;
        this.lastDroppedTime = 0;
    }

    /** Books one frame of `steps` steps, `elapsed` seconds handed in and `dropped` seconds lost. */
    record(steps, elapsed, dropped) {
        this.lastSteps = steps;
        this.lastElapsed = elapsed;
        this.lastDroppedTime = dropped;
        this.totalSteps += steps;
        this.simulatedTime += steps * this.stepDuration;
        this.droppedTime += dropped;
        return this.snapshot();
    }

    snapshot() {
        return {
            stepDuration: this.stepDuration,
            maxStepsPerFrame: this.maxStepsPerFrame,
            steps: this.lastSteps,
            elapsed: this.lastElapsed,
            totalSteps: this.totalSteps,
            simulatedTime: this.simulatedTime,
            backlogTime: this.accumulator,
            frameDroppedTime: this.lastDroppedTime,
            droppedTime: this.droppedTime,
            overloaded: this.lastDroppedTime > 0
        };
    }
}
