'use strict';
// Loads the N-API addon (p4a.node, linked against ../lib/libp4a.so).  There
// is no JavaScript fallback: a missing build is an error.
const path = require('path');
const fs = require('fs');

const file = path.join(__dirname, 'p4a.node');
if (!fs.existsSync(file)) {
  throw new Error(`${file} is missing: build it with __graft_entry__.build() ` +
                  '(particles4all_amd/build.py)');
}
module.exports = require(file);
