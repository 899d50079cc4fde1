'use strict';
// Renderer facade of the MI355X splat renderer (Node >= 12, CommonJS).
//
// Mirrors the reference Renderer's surface (src/renderer/renderer.js:160,
// 461-508): scenes are uploaded by name (uploadGsData) and selected (setGs);
// each render(camera) runs the whole pass graph on the GPU through the N-API
// addon and returns the image.  Per-pass GPU times are delivered to
// onProfile(rows) in the reference profiler's shape — { label: { ms } } plus
// TOTAL, averaged over up to `printEvery` frames and then reset
// (src/gpu/profiler.js:96-123) — from hipEvents instead of timestamp queries.
// Errors from the native side are thrown as Error, as the reference's fail()
// does (src/utils/utils.js:3-47).

const path = require('path');

let addon = null;
function native() {
    if (!addon) addon = require(path.join(__dirname, '..', 'build', 'gsr_napi.node'));
    return addon;
}

class Profiler {
    constructor(printEvery, onResult) {
        this.printEvery = printEvery;
        this.onResult = onResult;
        this.sums = new Map();
        this.order = [];
        this.frames = 0;
    }

    accumulate(pairs) {
        for (const [label, ms] of pairs) {
            if (label === 'TOTAL') continue;
            if (!this.sums.has(label)) this.order.push(label);
            this.sums.set(label, (this.sums.get(label) || 0) + ms);
        }
        ++this.frames;
        const rows = {};
        let total = 0;
        for (const label of this.order) {
            const avg = this.sums.get(label) / this.frames;
            total += avg;
            rows[label] = { ms: +avg.toFixed(3) };
        }
        rows.TOTAL = { ms: +total.toFixed(3) };
        if (this.onResult) this.onResult(rows);
        if (this.frames >= this.printEvery) {
            this.sums.clear();
            this.order = [];
            this.frames = 0;
        }
        return rows;
    }
}

class Renderer {
    // options: { device = 0, onProfile = null, printEvery = 8, tileSize = 16 }
    constructor(options) {
        const o = options || {};
        this.handle = native().create(o.device || 0);
        this.scenes = new Map();
        this.current = null;
        this.tileSize = o.tileSize || 16;
        this.profiler = o.onProfile ? new Profiler(o.printEvery || 8, o.onProfile) : null;
    }

    // data: { count, packed: Float32Array } as the PLY loader returns it.
    uploadGsData(data, name) {
        native().uploadGsData(this.handle, name, data.packed, data.count);
        this.scenes.set(name, data.count);
        if (this.current === null) this.current = name;
    }

    setGs(name) {
        if (!this.scenes.has(name)) throw new Error(`Invalid Gs buffer for name "${name}"`);
        native().setGs(this.handle, name);
        this.current = name;
    }

    freeGs(name) {
        native().freeGs(this.handle, name);
        this.scenes.delete(name);
        if (this.current === name) this.current = null;
    }

    // render(camera, { rgba8 = false, flipY = false })
    //   -> { width, height, data } with data a planar Float32Array [3][H][W]
    //      (or an RGBA8 Uint8Array [H][W][4] when rgba8 is set).
    render(camera, opts) {
        if (this.current === null) throw new Error('no scene uploaded');
        const o = opts || {};
        const count = this.scenes.get(this.current);
        const uniform = camera.uniformBlock(count);
        const data = native().render(this.handle, uniform, {
            tileSize: this.tileSize,
            timing: this.profiler !== null,
            rgba8: !!o.rgba8,
            flipY: !!o.flipY,
        });
        if (this.profiler) this.profiler.accumulate(native().timings(this.handle));
        return { width: camera.width, height: camera.height, data: data };
    }

    instanceCount() {
        return native().instanceCount(this.handle);
    }

    destroy() {
        if (this.handle) native().destroy(this.handle);
        this.handle = null;
    }
}

module.exports = { Renderer, Profiler, native };
