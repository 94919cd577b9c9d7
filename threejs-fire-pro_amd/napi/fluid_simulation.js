'use strict';
// The engine object Fire Pro's library layer uses (src/library/FireSimulation.ts:203-209,
// :211, :296-333, :358, :372, :457-459, :517, :522, :539), backed by the MI355X solver
// through the N-API addon instead of WebGPU passes. Members keep the reference's names
// and meaning (src/engine/FluidSimulation.ts); frames are the reference's plain
// SimulationFrame objects (src/engine/types.ts:235-311).
const path = require('path');

const addon = require(path.join(__dirname, 'firepro.node'));

class FluidSimulation {
  constructor(grid) {
    this.handle = addon.create(grid);
    this.voxelSize = grid.voxelSize;
    this.brickCells = grid.brickSize === undefined ? 16 : grid.brickSize;
    this.velocityDivisor = grid.velocityDivisor === undefined ? 1 : grid.velocityDivisor;
    this.smokeDivisor = grid.smokeDivisor === undefined ? 1 : grid.smokeDivisor;
    this.steps = 0;
    this.cutoffValue = 0;
    this.budget = Infinity;
    this.groundValue = true;
  }

  /** FluidSimulation.create (FluidSimulation.ts:419-422); the device is the addon's GPU. */
  static async create(_device, grid) {
    return new FluidSimulation(grid);
  }

  /** step (FluidSimulation.ts:1367-1373): enqueues one substep; returns once its activity
   * readback has landed, the rest of the step still running on the device. */
  step(frame, dt = 1 / 60) {
    addon.step(this.handle, frame, dt);
    this.steps++;
  }

  reset() {
    addon.reset(this.handle);
    this.steps = 0;
  }

  /** dispose (FluidSimulation.ts:1780-1802): idempotent. */
  dispose() {
    addon.dispose(this.handle);
  }

  get cutoff() { return this.cutoffValue; }
  set cutoff(value) {
    addon.setCutoff(this.handle, value);
    this.cutoffValue = value;
  }
  get voxelBudget() { return this.budget; }
  set voxelBudget(value) {
    addon.setVoxelBudget(this.handle, value);
    this.budget = value;
  }
  get ground() { return this.groundValue; }
  set ground(value) {
    addon.setGround(this.handle, Boolean(value));
    this.groundValue = Boolean(value);
  }

  get time() { return addon.stats(this.handle).time; }
  get activeBrickCount() { return addon.stats(this.handle).activeBrickCount; }
  get poolSlots() { return addon.stats(this.handle).poolSlots; }
  get poolLimited() { return addon.stats(this.handle).poolLimited; }
  get memoryBytes() { return addon.stats(this.handle).memoryBytes; }
  get meshMemoryBytes() { return addon.stats(this.handle).meshMemoryBytes; }
  get forceMemoryBytes() { return addon.stats(this.handle).forceMemoryBytes; }
  /** renderBricks (FluidSimulation.ts:781-795): the boxes and tile words the renderer uses. */
  get renderBricks() { return addon.renderBricks(this.handle); }

  /** showSolverField (FluidSimulation.ts:1542-1549). */
  showSolverField(field) {
    addon.showSolverField(this.handle, field);
  }
  /** The shown solver field per active brick: {ids, values}. */
  readSolverField() {
    return addon.readSolverField(this.handle);
  }

  /** readBricks (FluidSimulation.ts:1749-1779): the bricks with a slot, in world space. */
  async readBricks() {
    const mins = addon.readBricks(this.handle);
    const size = this.brickCells * this.voxelSize;
    const out = [];
    for (let i = 0; i < mins.length; i += 3) {
      const min = [mins[i], mins[i + 1], mins[i + 2]];
      out.push({ min, max: min.map((v) => v + size) });
    }
    return out;
  }

  /** The multigrid pressure policy (no reference counterpart: it runs one V-cycle). */
  setPressure(mode, tolerance = 1e-6, maxCycles = 20) {
    addon.setPressure(this.handle, mode, tolerance, maxCycles);
  }

  /** The headless ray-march: width x height premultiplied (radiance, 1 - T). */
  render(camera, settings) {
    return addon.render(this.handle, camera, settings);
  }
}

module.exports = { FluidSimulation, version: addon.version };
