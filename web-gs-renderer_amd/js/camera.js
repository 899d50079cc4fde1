'use strict';
// Camera and the 176-byte per-frame uniform block (Node >= 12, CommonJS).
//
// Conventions of the reference camera (src/renderer/camera.js:87-98,
// 135-159) with wgpu-matrix semantics: view = lookAt(eye, eye + forward, up)
// looking down -z, projection = perspective(fovY, W/H, 0.2, 1000) with clip
// z in [0, 1], focal = (W / (2 tanFovX), H / (2 tanFovY)), tanFovX =
// tanFovY * W / H.  Matrices are column-major Float32Arrays computed in
// doubles.  Keyboard/mouse fly controls (camera.js:38-50, 100-133) are
// browser input and are not part of this host.

const Z_NEAR = 0.2;
const Z_FAR = 1000.0;
const WORLD_UP = [0, 1, 0];

function sub(a, b) { return [a[0] - b[0], a[1] - b[1], a[2] - b[2]]; }
function add(a, b) { return [a[0] + b[0], a[1] + b[1], a[2] + b[2]]; }
function dot(a, b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
function cross(a, b) { return [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]]; }
function normalize(a) { const l = Math.sqrt(dot(a, a)); return l > 0 ? [a[0] / l, a[1] / l, a[2] / l] : [0, 0, 0]; }
function degToRad(d) { return d * Math.PI / 180; }

function lookAt(eye, target, up) {
    const z = normalize(sub(eye, target));
    const x = normalize(cross(up, z));
    const y = normalize(cross(z, x));
    const m = new Float32Array(16);
    m[0] = x[0]; m[1] = y[0]; m[2] = z[0]; m[3] = 0;
    m[4] = x[1]; m[5] = y[1]; m[6] = z[1]; m[7] = 0;
    m[8] = x[2]; m[9] = y[2]; m[10] = z[2]; m[11] = 0;
    m[12] = -dot(x, eye); m[13] = -dot(y, eye); m[14] = -dot(z, eye); m[15] = 1;
    return m;
}

function perspective(fovy, aspect, near, far) {
    const f = Math.tan(Math.PI * 0.5 - 0.5 * fovy);
    const m = new Float32Array(16);
    m[0] = f / aspect;
    m[5] = f;
    m[11] = -1;
    const rangeInv = 1 / (near - far);
    m[10] = far * rangeInv;
    m[14] = far * near * rangeInv;
    return m;
}

function multiply(a, b) {   // column-major a * b
    const m = new Float32Array(16);
    for (let c = 0; c < 4; ++c) {
        for (let r = 0; r < 4; ++r) {
            let s = 0;
            for (let k = 0; k < 4; ++k) s += a[k * 4 + r] * b[c * 4 + k];
            m[c * 4 + r] = s;
        }
    }
    return m;
}

class Camera {
    // Either a reference-style yaw/pitch camera (defaults of camera.js:5-14,
    // 57-72) or, with opts.target, a look-at camera.
    constructor(width, height, opts) {
        const o = opts || {};
        this.width = width;
        this.height = height;
        this.fov = o.fov !== undefined ? o.fov : 60.0;
        this.eye = o.eye ? o.eye.slice() : [0, -2, -10];
        this.yaw = o.yaw !== undefined ? o.yaw : 90.0;
        this.pitch = o.pitch !== undefined ? o.pitch : 0.0;
        this.target = o.target ? o.target.slice() : null;
        this.up = o.up ? o.up.slice() : WORLD_UP.slice();
    }

    resize(width, height) {
        this.width = width;
        this.height = height;
    }

    frame() {
        const yaw = degToRad(this.yaw);
        const pitch = degToRad(this.pitch);
        const forward = normalize([Math.cos(yaw) * Math.cos(pitch), Math.sin(pitch), Math.sin(yaw) * Math.cos(pitch)]);
        const right = normalize(cross(forward, WORLD_UP));
        const up = normalize(cross(right, forward));
        return { forward: forward, right: right, up: up };
    }

    viewMatrix() {
        if (this.target) return lookAt(this.eye, this.target, this.up);
        const f = this.frame();
        return lookAt(this.eye, add(this.eye, f.forward), f.up);
    }

    // { view, projView, eye, focal, tanFov } as camera.js:87-98 returns it.
    getUniform() {
        const ratio = this.width / this.height;
        const tanFovY = Math.tan(degToRad(this.fov) / 2);
        const tanFovX = tanFovY * ratio;
        const view = this.viewMatrix();
        const proj = perspective(degToRad(this.fov), ratio, Z_NEAR, Z_FAR);
        return {
            view: view,
            projView: multiply(proj, view),
            eye: this.eye.slice(),
            focal: [this.width / (2 * tanFovX), this.height / (2 * tanFovY)],
            tanFov: [tanFovX, tanFovY],
        };
    }

    // The GlobalUniforms block packed as renderer.js:773-797 does: view @0,
    // viewProj @16, eye @32, count (u32) @35, focal @36, tanFov @38,
    // textureSize @40, two padding words.
    uniformBlock(count) {
        const u = this.getUniform();
        const bytes = new ArrayBuffer(176);
        const f = new Float32Array(bytes);
        const i = new Uint32Array(bytes);
        f.set(u.view, 0);
        f.set(u.projView, 16);
        f.set(u.eye, 32);
        i[35] = count >>> 0;
        f.set(u.focal, 36);
        f.set(u.tanFov, 38);
        f.set([this.width, this.height], 40);
        return f;
    }
}

module.exports = { Camera, lookAt, perspective, multiply, Z_NEAR, Z_FAR };
