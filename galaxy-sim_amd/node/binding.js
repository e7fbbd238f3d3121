// Loads the N-API addon (galaxy_hip.node, built by galaxy_hip/build.py) that
// binds libgalaxy_hip.so.  Shared by NBodyGPU.js and mortonTree.js.
import { createRequire } from 'module';

const require = createRequire(import.meta.url);
export const addon = require('./galaxy_hip.node');
