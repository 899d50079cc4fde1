'use strict';
// Node entry point of the MI355X Gaussian-splat renderer's JS host.
const { Renderer, Profiler } = require('./renderer');
const { Camera, lookAt, perspective } = require('./camera');
const { parsePly, loadPlyFile } = require('./loader');

module.exports = { Renderer, Profiler, Camera, lookAt, perspective, parsePly, loadPlyFile };
