'use strict';
// PLY loader for Gaussian-splat scenes (Node >= 12, CommonJS).
//
// Behaviour follows the reference loader (src/loader/loader.js:1-213):
// binary little-endian only; the vertex element's 59 required properties
// (x y z, scale_0..2, rot_0..3, opacity, f_dc_0..2, f_rest_0..44) in any order
// and any of the 16 scalar types; list properties are rejected; the output is
// one Float32Array in the block layout pos[3N] | scale[3N] | rot[4N] |
// f_dc[3N] | opacity[N] | f_rest[45N] (loader.js:75-85).  Failures throw an
// Error with a message naming the file, like the reference's fail()
// (src/utils/utils.js:3-47).  Unlike the reference, the header end is found
// without decoding the whole body as text, and a body shorter than
// count * stride is reported instead of surfacing as a DataView RangeError.

const REQUIRED = ['x', 'y', 'z', 'scale_0', 'scale_1', 'scale_2', 'rot_0', 'rot_1', 'rot_2', 'rot_3',
    'opacity', 'f_dc_0', 'f_dc_1', 'f_dc_2'];
for (let i = 0; i < 45; ++i) REQUIRED.push('f_rest_' + i);

const TYPE_SIZE = {
    char: 1, uchar: 1, int8: 1, uint8: 1,
    short: 2, ushort: 2, int16: 2, uint16: 2,
    int: 4, uint: 4, int32: 4, uint32: 4,
    float: 4, float32: 4,
    double: 8, float64: 8,
};

const READ = {
    char: (v, o) => v.getInt8(o), int8: (v, o) => v.getInt8(o),
    uchar: (v, o) => v.getUint8(o), uint8: (v, o) => v.getUint8(o),
    short: (v, o) => v.getInt16(o, true), int16: (v, o) => v.getInt16(o, true),
    ushort: (v, o) => v.getUint16(o, true), uint16: (v, o) => v.getUint16(o, true),
    int: (v, o) => v.getInt32(o, true), int32: (v, o) => v.getInt32(o, true),
    uint: (v, o) => v.getUint32(o, true), uint32: (v, o) => v.getUint32(o, true),
    float: (v, o) => v.getFloat32(o, true), float32: (v, o) => v.getFloat32(o, true),
    double: (v, o) => v.getFloat64(o, true), float64: (v, o) => v.getFloat64(o, true),
};

function fail(message) {
    throw new Error(message);
}

// Byte offset just past "end_header" and its line break.
function headerEnd(bytes, fileName) {
    const tag = Buffer.from('end_header', 'ascii');
    const idx = Buffer.from(bytes.buffer, bytes.byteOffset, bytes.byteLength).indexOf(tag);
    if (idx < 0) fail(`Invalid PLY (${fileName}): "end_header" not found`);
    let end = idx + tag.length;
    if (bytes[end] === 13) ++end;
    if (bytes[end] === 10) ++end;
    return end;
}

function parseHeader(fileName, text) {
    let count = 0;
    let hasFormat = false;
    let littleEndian = false;
    let element = null;
    const properties = [];
    let stride = 0;
    for (const raw of text.split('\n')) {
        const w = raw.trim().split(/\s+/);
        if (w[0] === 'format') {
            hasFormat = true;
            littleEndian = w[1] === 'binary_little_endian';
        } else if (w[0] === 'element') {
            element = w[1];
            if (element === 'vertex') count = parseInt(w[2], 10);
        } else if (w[0] === 'property' && element === 'vertex') {
            if (w[1] === 'list') fail(`Invalid PLY (${fileName}): list properties are not supported`);
            const type = w[1];
            if (!(type in TYPE_SIZE)) fail(`Invalid PLY (${fileName}): unknown property type "${type}"`);
            properties.push({ name: w[2], type: type, offset: stride });
            stride += TYPE_SIZE[type];
        }
    }
    if (!hasFormat) fail(`Invalid PLY (${fileName}): missing "format" line`);
    if (!littleEndian) fail(`Invalid PLY (${fileName}): unsupported format (binary_little_endian only)`);
    if (!(count > 0)) fail(`Invalid PLY (${fileName}): vertex count missing or zero`);
    return { count: count, properties: properties, stride: stride };
}

// parsePly(fileName, ArrayBuffer | Buffer | Uint8Array) -> { count, packed: Float32Array }
function parsePly(fileName, data) {
    const bytes = data instanceof Uint8Array ? data : new Uint8Array(data);
    const body = headerEnd(bytes, fileName);
    const text = Buffer.from(bytes.buffer, bytes.byteOffset, body).toString('ascii');
    const header = parseHeader(fileName, text);
    const byName = new Map(header.properties.map((p) => [p.name, p]));
    for (const name of REQUIRED) {
        if (!byName.has(name)) fail(`Invalid PLY (${fileName}): missing required property "${name}"`);
    }
    const n = header.count;
    const stride = header.stride;
    if (body + n * stride > bytes.byteLength) {
        fail(`Invalid PLY (${fileName}): body holds ${bytes.byteLength - body} bytes, ${n} vertices need ${n * stride}`);
    }
    const view = new DataView(bytes.buffer, bytes.byteOffset + body, n * stride);
    const packed = new Float32Array(n * 59);
    // destination of each required property: [block base, per-splat width, slot]
    const dest = [];
    const blocks = [[0, 3, ['x', 'y', 'z']], [3 * n, 3, ['scale_0', 'scale_1', 'scale_2']],
        [6 * n, 4, ['rot_0', 'rot_1', 'rot_2', 'rot_3']], [10 * n, 3, ['f_dc_0', 'f_dc_1', 'f_dc_2']],
        [13 * n, 1, ['opacity']], [14 * n, 45, REQUIRED.slice(14)]];
    for (const [base, width, names] of blocks) {
        names.forEach((name, slot) => {
            const p = byName.get(name);
            dest.push({ base: base, width: width, slot: slot, offset: p.offset, read: READ[p.type] });
        });
    }
    for (let i = 0; i < n; ++i) {
        const row = i * stride;
        for (let k = 0; k < dest.length; ++k) {
            const d = dest[k];
            packed[d.base + i * d.width + d.slot] = d.read(view, row + d.offset);
        }
    }
    return { count: n, packed: packed };
}

function loadPlyFile(path) {
    const fs = require('fs');
    return parsePly(require('path').basename(path), fs.readFileSync(path));
}

module.exports = { parsePly, loadPlyFile, REQUIRED, TYPE_SIZE };
