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

// parseIni (settingsini.js:45-55)
function parseIni(text) {
  const kv = {};
  for (const raw of String(text).split(/\r?\n/)) {
    const line = raw.trim();
    if (!line || line.startsWith('#') || line.startsWith(';')) continue;
    const eq = line.indexOf('=');
    if (eq <= 0) continue;
    kv[line.slice(0, eq).trim().toLowerCase()] = line.slice(eq + 1).trim();
  }
  return kv;
}

const VIEW = { 0: 'particles', 1: 'mesh', 2: 'mesh', 3: 'ray', 4: 'ssfr' };

// iniToQuery (settingsini.js:59-93)
function iniToQuery(text, opts) {
  opts = opts || {};
  const kv = parseIni(text);
  const sp = new URLSearchParams();
  const applied = [], skipped = [];
  for (const k of Object.keys(kv)) {
    const v = kv[k];
    const q = KEYS[k];
    if (q) { sp.set(q, v); applied.push(k); continue; }
    if (k === 'display') {
      const view = VIEW[String(parseInt(v, 10))];
      if (view) { sp.set('view', view); applied.push(k); } else skipped.push(k);
      continue;
    }
    if (k === 'cubemap') continue;
    if (k === 'body') {
      if (v.trim()) sp.set('body', v.trim()); else sp.set('bodies', '0');
      applied.push(k);
      continue;
    }
    skipped.push(k);
  }
  let missingPanorama = null;
  if (kv.cubemap) {
    const name = kv.cubemap.replace(/^"|"$/g, '').split(/[\\/]/).pop();
    if (name) {
      sp.set('cubemap', 'env/' + name);
      applied.push('cubemap');
      missingPanorama = name;
    }
  }
  const keep = opts.keep || {};
  for (const k of Object.keys(keep)) {
    const v = keep[k];
    if (v !== null && v !== undefined && !sp.has(k)) sp.set(k, v);
  }
  return { query: sp.toString(), applied, skipped, missingPanorama };
}

// The simulation part of main.js:361-390: URL query -> params.
function paramsFromQuery(query, preset) {
  const q = new URLSearchParams(query);
  const qNum = (k, d) => (q.has(k) ? Number(q.get(k)) : d);
  const qVec = (k, n) => {
    const v = q.get(k);
    if (!v) return null;
    const a = v.split(/[ ,]+/).map(Number);
    return a.length >= n && a.every(x => Number.isFinite(x)) ? a : null;
  };
  const params = defaultParams(preset);
  const bodiesOn = q.get('bodies') !== '0';
  if (!bodiesOn) params.bodies = [];
  params.targetParticleCount = qNum('particles', params.targetParticleCount);
  if (bodiesOn && q.get('body')) params.bodies = q.get('body').split(',');
  params.bodySize = qNum('bodysize', params.bodySize);
  params.substeps = qNum('substeps', params.substeps);
  params.iterations = qNum('iters', params.iterations);
  params.omega = qNum('omega', params.omega);
  params.xsphC = qNum('xsph', params.xsphC);
  params.sCorrK = qNum('scorr', params.sCorrK);
  params.spacing = qNum('spacing', params.spacing);
  params.gravity = qNum('gravity', params.gravity);
  params.surfaceTensionK = qNum('tension', params.surfaceTensionK);
  params.cfmEpsilonRel = qNum('cfm', params.cfmEpsilonRel);
  params.sCorrDq = qNum('scorrdq', params.sCorrDq);
  params.noBoundary = q.get('boundary') === '0';
  const b = qVec('box', 3);
  if (b) params.box = b.slice(0, 3);
  return params;
}

// Construction from a preset: presets/<name>.ini through iniToQuery, URL
// overrides first (main.js:255-279), then paramsFromQuery.
function loadPreset(name, overrides, dir) {
  const preset = PRESETS[name] ? name : 'small';
  const q = new URLSearchParams(overrides || '');
  let note = null;
  if (PRESETS[preset].ini) {
    const file = path.join(dir || path.join(__dirname, '..'), PRESETS[preset].ini);
    const r = iniToQuery(fs.readFileSync(file, 'utf8'));
    for (const kv of new URLSearchParams(r.query)) if (!q.has(kv[0])) q.set(kv[0], kv[1]);
    const unknown = r.skipped.filter(k => UNSUPPORTED.indexOf(k) < 0);
    if (unknown.length) note = `${file}: ${unknown.length} unknown keys (${unknown.join(', ')})`;
  }
  const params = paramsFromQuery(q.toString(), preset);
  return { params, query: q.toString(), note };
}

module.exports = { KEYS, UNSUPPORTED, parseIni, iniToQuery, paramsFromQuery, loadPreset };
