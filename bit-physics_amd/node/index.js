'use strict';
// Loader of the Node-API addon over libd3q19 (types in index.d.ts).
const path = require('path');
const addon = require(path.join(__dirname, 'build', 'd3q19.node'));

const FACES = { '-x': 0, '+x': 1, '-y': 2, '+y': 3, '-z': 4, '+z': 5 };

class Lattice extends addon.Lattice {
  constructor(cfg) {
    const c = Object.assign({}, cfg);
    if (Array.isArray(c.wallFaces)) {
      c.wallFaces = c.wallFaces.reduce((m, f) => m | (1 << FACES[f]), 0);
    }
    super(c);
    this.shape = [c.nx, c.ny, c.nz];
  }
}

module.exports = { Lattice, version: addon.version };
