% Build the MATLAB gateways over libmcsart_hip (run from this folder after
% `python -m mcsart_amd.build` has produced ../libmcsart_hip.so).
lib = fullfile('..');
inc = fullfile('..', '..', 'include');
for g = {'entry_AUMISART', 'mex_backprojection', 'mex_projection', 'mex_diff'}
    mex('-R2018a', ['-I' inc], ['-L' lib], '-lmcsart_hip', ['LDFLAGS=$LDFLAGS -Wl,-rpath,' lib], ...
        [g{1} '.cpp']);
end
