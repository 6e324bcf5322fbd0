// Minimal assert-based runner for the JS host tests (Node 12 has no test
// runner): `node tests/js/run.mjs <file.test.mjs>...`; each module exports
// `tests`, an object of named (possibly async) functions.
import path from 'path';
import { pathToFileURL } from 'url';

async function main() {
    let failed = 0;
    let passed = 0;
    for (const file of process.argv.slice(2)) {
        const mod = await import(pathToFileURL(path.resolve(file)).href);
        for (const [name, fn] of Object.entries(mod.tests)) {
            try {
                await fn();
                passed++;
                console.log(`ok   ${path.basename(file)} :: ${name}`);
            } catch (error) {
                failed++;
                console.log(`FAIL ${path.basename(file)} :: ${name}\n${error && error.stack}`);
            }
        }
    }
    console.log(`${passed} passed, ${failed} failed`);
    process.exit(failed ? 1 : 0);
}

main().catch((error) => {
    console.error(error);
    process.exit(1);
});
