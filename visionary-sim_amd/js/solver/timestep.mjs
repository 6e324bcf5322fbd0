// The fixed physical timestep (mirrors the semantics of
// packages/visionary-sim/src/sim/solver/timestep.js:17-180): wall time is
// banked, simulated in whole fixed steps, and whole steps past the per-frame
// budget are dropped and counted rather than run later or enlarged.

export const TARGET_CFL = 0.75;
export const MAX_STEP_DURATION = 1 / 240;
export const MAX_STEPS_PER_FRAME = 24;

export function stableStepDuration(waveRate) {
    if (!Number.isFinite(waveRate) || waveRate < 0) {
        throw new RangeError(`waveRate must be a finite non-negative number, got ${waveRate}`);
    }
    return waveRate === 0 ? MAX_STEP_DURATION : Math.min(MAX_STEP_DURATION, TARGET_CFL / waveRate);
}

function checkDuration(value) {
    if (!Number.isFinite(value) || value <= 0) {
        throw new RangeError(`stepDuration must be a finite positive number, got ${value}`);
    }
}

export class FixedTimestep {
    constructor(stepDuration, options = {}) {
        const maxStepsPerFrame = options.maxStepsPerFrame === undefined ? MAX_STEPS_PER_FRAME : options.maxStepsPerFrame;
        checkDuration(stepDuration);
        if (!Number.isInteger(maxStepsPerFrame) || maxStepsPerFrame < 1) {
            throw new RangeError(`maxStepsPerFrame must be a positive integer, got ${maxStepsPerFrame}`);
        }
        this.stepDuration = stepDuration;
        this.maxStepsPerFrame = maxStepsPerFrame;
        this.reset();
    }

    setStepDuration(stepDuration) {
        checkDuration(stepDuration);
        this.stepDuration = stepDuration;
    }

    consume(elapsed, options = {}) {
        if (!Number.isFinite(elapsed) || elapsed < 0) {
            throw new RangeError(`elapsed must be a finite non-negative number, got ${elapsed}`);
        }
        this.accumulator += elapsed;
        const slack = this.stepDuration * 1e-9;
        const due = Math.floor((this.accumulator + slack) / this.stepDuration);
        let steps = Math.min(due, this.maxStepsPerFrame);
        const dropped = due - steps;
        this.accumulator -= due * this.stepDuration;
        if (this.accumulator < 0 && this.accumulator > -slack) this.accumulator = 0;
        if (options.forceStep && steps === 0) steps = 1;
        this.lastSteps = steps;
        this.lastElapsed = elapsed;
        this.lastDroppedTime = dropped * this.stepDuration;
        this.simulatedTime += steps * this.stepDuration;
        this.totalSteps += steps;
        this.droppedTime += this.lastDroppedTime;
        return this.snapshot();
    }

    idle() {
        this.lastSteps = 0;
        this.lastElapsed = 0;
        this.lastDroppedTime = 0;
        return this.snapshot();
    }

    singleStep() {
        this.lastSteps = 1;
        this.lastElapsed = 0;
        this.lastDroppedTime = 0;
        this.simulatedTime += this.stepDuration;
        this.totalSteps++;
        return this.snapshot();
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
