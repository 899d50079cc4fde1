'use strict';
// Test driver for the Node host: node tests/js_driver.js <command> <json-args>
// Writes results as JSON on stdout (binary payloads go to files named in args).
const fs = require('fs');
const path = require('path');
const host = require(path.join(__dirname, '..', 'web-gs-renderer_amd', 'js', 'index.js'));

function main() {
    const cmd = process.argv[2];
    const args = JSON.parse(process.argv[3] || '{}');
    if (cmd === 'parse') {
        const r = host.loadPlyFile(args.ply);
        fs.writeFileSync(args.out, Buffer.from(r.packed.buffer, r.packed.byteOffset, r.packed.byteLength));
        return { count: r.count };
    }
    if (cmd === 'uniform') {
        const cam = new host.Camera(args.width, args.height, { eye: args.eye, target: args.target, up: args.up, fov: args.fov });
        const u = cam.uniformBlock(args.count);
        fs.writeFileSync(args.out, Buffer.from(u.buffer));
        return { ok: true };
    }
    if (cmd === 'profiler') {
        const seen = [];
        const p = new host.Profiler(args.printEvery, (rows) => seen.push(rows));
        for (const frame of args.frames) p.accumulate(frame);
        return { rows: seen };
    }
    if (cmd === 'create') {
        const r = new host.Renderer({ device: 0 });
        r.destroy();
        return { ok: true };
    }
    if (cmd === 'render') {
        const profiles = [];
        const r = new host.Renderer({ device: 0, onProfile: (rows) => profiles.push(rows) });
        const scene = host.loadPlyFile(args.ply);
        r.uploadGsData(scene, 'scene');
        r.setGs('scene');
        const cam = new host.Camera(args.width, args.height, { eye: args.eye, target: args.target, up: args.up, fov: args.fov });
        const img = r.render(cam);
        fs.writeFileSync(args.out, Buffer.from(img.data.buffer, img.data.byteOffset, img.data.byteLength));
        const rgba = r.render(cam, { rgba8: true, flipY: true });
        fs.writeFileSync(args.out_rgba, Buffer.from(rgba.data.buffer, rgba.data.byteOffset, rgba.data.byteLength));
        let bad = null;
        try { r.setGs('missing'); } catch (e) { bad = e.message; }
        const instances = r.instanceCount();
        r.destroy();
        return { instances: instances, profiles: profiles.length, lastProfile: profiles[profiles.length - 1], badScene: bad };
    }
    throw new Error('unknown command ' + cmd);
}

try {
    process.stdout.write(JSON.stringify({ ok: true, result: main() }));
} catch (e) {
    process.stdout.write(JSON.stringify({ ok: false, error: e.message }));
}
