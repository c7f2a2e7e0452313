'use strict';
// winding_amd: the splat path's JavaScript API (CommonJS; index.mjs re-exports it as an ES module).
const splats = require('./splats.js');
const math = require('./math.js');
const learned = require('./learned.js');

module.exports = Object.assign({}, splats, math, learned);
