'use strict';
// Preset loading for the JS host: the reference's settings reader
// (src/settingsini.js:3-93) and the query -> simulation-parameter plumbing
// of its app shell (src/main.js:255-279, 361-390), reduced to the keys that
// reach Sim.  Rendering and UI keys are parsed and reported as applied or
// skipped exactly as the reference does, then ignored.

const fs = require('fs');
const path = require('path');
const { defaultParams, PRESETS } = require('./scene');

// pbf_settings.ini key -> URL key (settingsini.js:3-33)
const KEYS = {
  substeps: 'substeps', iterations: 'iters', cfm: 'cfm', scorrk: 'scorr',
  scorrdq: 'scorrdq', xsph: 'xsph', tension: 'tension', gravity: 'gravity',
  omega: 'omega', particles: 'particles', spacing: 'spacing',
  meshres: 'meshres', meshiso: 'meshiso', fieldsmooth: 'fieldsmooth',
  normalsmooth: 'normalsmooth', raysurface: 'raysurface',
  anisoratio: 'anisoratio', anisolambda: 'anisolambda', anisonbrs: 'anisonbrs',
  anisolonely: 'anisolonely', anisoradius: 'anisoradius',
  anisostretch: 'anisostretch', anisoks: 'anisoks',
  ior: 'ior', absorption: 'absorption', thicksteps: 'thicksteps',
  groundreflection: 'groundrefl', roughness: 'roughness',
  sunintensity: 'sunint', sunelev: 'sunelev', sunazim: 'sunazim',
  exposure: 'exposure',
  ssfrradius: 'ssfrradius', ssfrfilter: 'ssfrfilter', ssfriters: 'ssfriters',
  ssfrsigma: 'ssfrsigma', ssfrdelta: 'ssfrdelta', ssfrmu: 'ssfrmu',
  ssfrrange: 'ssfrrange', ssfrcleanup: 'ssfrcleanup', ssfrthick: 'ssfrthick',
  ssfrthickr: 'ssfrthickr', ssfrthickblur: 'ssfrthickblur',
  ssfrdepthcull: 'ssfrdepthcull', ssfrscale: 'ssfrscale',
  renderradius: 'radius', timescale: 'timescale', colorspeedmax: 'speedmax',
  hover: 'force', hoverradius: 'fradius', hoverstrength: 'fstrength',
  hoverlimit: 'flimit', grabstrength: 'grabstrength',
  pourspeed: 'pourspeed', pourwidth: 'pourwidth', pourheight: 'pourheight',
  pourtilt: 'pourtilt',
  floorplane: 'floorplane', cubemapintensity: 'cubemapintensity',
  cubemapyaw: 'cubemapyaw',
  box: 'box', transmit: 'transmit', camera: 'camera',
  bodysize: 'bodysize',
};

// keys the reference reads but does not support (settingsini.js:35-43)
const UNSUPPORTED = [
  'showkernels', 'kernelscale', 'soraverage', 'fieldfilter',
  'ssfrstretch', 'ssfrstretchmax', 'ssfrstretchthin', 'ssfrstretchreach',
  'ssfrhalfthick', 'weburl', 'webhighperf', 'webpresets',
];

// Reader of the reference's preset format (behaviour of settingsini.js:45-55):
// one `key = value` entry per line, keys case-insensitive, `#` and `;`
// comment lines, a line without a key before its first `=` ignored, the
// last entry of a key winning.
const ENTRY = /^([^#;=][^=]*)=(.*)$/;

function parseIni(text) {
  const out = {};
  String(text).split(/\r?\n/).forEach((raw) => {
    const m = ENTRY.exec(raw.trim());
    if (m) out[m[1].trim().toLowerCase()] = m[2].trim();
  });
  return out;
}

const VIEW = { 0: 'particles', 1: 'mesh', 2: 'mesh', 3: 'ray', 4: 'ssfr' };
const own = (o, k) => Object.prototype.hasOwnProperty.call(o, k);

// The URL entries one ini key contributes, or null for a key or value the
// app does not understand.
function urlEntries(key, value) {
  if (own(KEYS, key)) return [[KEYS[key], value]];
  if (key === 'display') {
    const view = VIEW[parseInt(value, 10)];
    return view ? [['view', view]] : null;
  }
  if (key === 'body') {
    const list = value.trim();
    return list ? [['body', list]] : [['bodies', '0']];
  }
  return null;
}

// File name of a panorama path, without the quotes a path may carry.
function panoramaName(value) {
  let v = value;
  if (v.startsWith('"')) v = v.slice(1);
  if (v.endsWith('"')) v = v.slice(0, -1);
  return v.slice(Math.max(v.lastIndexOf('/'), v.lastIndexOf('\\')) + 1);
}

// A preset file as the URL query the app starts from (behaviour of
// settingsini.js:59-93): the keys it took (`applied`), the ones it did not
// (`skipped`), and the panorama a server has to provide.  `opts.keep` adds
// entries the file does not set.
function iniToQuery(text, opts) {
  const kv = parseIni(text);
  const entries = [];
  const applied = [];
  const skipped = [];
  Object.keys(kv).filter(k => k !== 'cubemap').forEach((k) => {
    const got = urlEntries(k, kv[k]);
    (got ? applied : skipped).push(k);
    if (got) got.forEach(e => entries.push(e));
  });
  const panorama = kv.cubemap ? panoramaName(kv.cubemap) : '';
  if (panorama) {
    entries.push(['cubemap', 'env/' + panorama]);
    applied.push('cubemap');
  }
  const sp = new URLSearchParams();
  entries.forEach(e => sp.set(e[0], e[1]));
  const keep = (opts && opts.keep) || {};
  Object.keys(keep).forEach((k) => {
    if (keep[k] !== null && keep[k] !== undefined && !sp.has(k)) sp.set(k, keep[k]);
  });
  return { query: sp.toString(), applied, skipped, missingPanorama: panorama || null };
}

// URL key -> Sim parameter, for the numeric parameters the app reads from
// the query (main.js:361-390).
const NUMERIC = [
  ['particles', 'targetParticleCount'], ['bodysize', 'bodySize'], ['substeps', 'substeps'],
  ['iters', 'iterations'], ['omega', 'omega'], ['xsph', 'xsphC'], ['scorr', 'sCorrK'],
  ['spacing', 'spacing'], ['gravity', 'gravity'], ['tension', 'surfaceTensionK'],
  ['cfm', 'cfmEpsilonRel'], ['scorrdq', 'sCorrDq'],
];

// Three or more finite numbers separated by spaces or commas, else null.
function numberList(text, atLeast) {
  if (!text) return null;
  const nums = text.split(/[ ,]+/).map(Number);
  if (nums.length < atLeast || nums.some(x => !Number.isFinite(x))) return null;
  return nums;
}

// The simulation parameters a URL query selects on top of a preset's
// defaults: numbers (a present key is taken as Number(value)), the body
// list (`bodies=0` empties it, `body=a,b` replaces it), `boundary=0`, and a
// box of three numbers.
function paramsFromQuery(query, preset) {
  const q = new URLSearchParams(query);
  const params = defaultParams(preset);
  NUMERIC.forEach(([key, name]) => {
    if (q.has(key)) params[name] = Number(q.get(key));
  });
  if (q.get('bodies') === '0') params.bodies = [];
  else if (q.get('body')) params.bodies = q.get('body').split(',');
  params.noBoundary = q.get('boundary') === '0';
  const box = numberList(q.get('box'), 3);
  if (box) params.box = box.slice(0, 3);
  return params;
}

// Construction from a preset: presets/<name>.ini read into a query, the
// caller's overrides winning over the file (main.js:255-279), then the
// parameters of that query.
function loadPreset(name, overrides, dir) {
  const preset = PRESETS[name] ? name : 'small';
  const q = new URLSearchParams(overrides || '');
  let note = null;
  if (PRESETS[preset].ini) {
    const file = path.join(dir || path.join(__dirname, '..'), PRESETS[preset].ini);
    const r = iniToQuery(fs.readFileSync(file, 'utf8'));
    new URLSearchParams(r.query).forEach((v, k) => { if (!q.has(k)) q.set(k, v); });
    const unknown = r.skipped.filter(k => UNSUPPORTED.indexOf(k) < 0);
    if (unknown.length) note = `${file}: ${unknown.length} unknown keys (${unknown.join(', ')})`;
  }
  return { params: paramsFromQuery(q.toString(), preset), query: q.toString(), note };
}

module.exports = { KEYS, UNSUPPORTED, parseIni, iniToQuery, paramsFromQuery, loadPreset };
