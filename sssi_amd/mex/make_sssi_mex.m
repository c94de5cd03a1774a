% make_sssi_mex  Build the MATLAB gateways of libsssi (run from this folder).
%
% Each gateway keeps the name and signature of the MATLAB/C-MEX function it
% replaces (src/fd_cmex/make.m builds the reference's CPU versions) and calls
% the C ABI declared in include/sssi.h.  libsssi.so must be built first
% (python -c "import sssi_amd.build as b; b.build()").
here = fileparts(mfilename('fullpath'));
inc = fullfile(here, '..', '..', 'include');
libdir = fullfile(here, '..');
gateways = {'fwdTimeCpmlFor2dAw_mex', 'rvsTimeCpmlFor2dAw_mex', ...
            'fwdTimeCpmlFor3dAw_mex', 'fwdTimeSpongeFor2dAw_mex', ...
            'rvsTimeSpongeFor2dAw_mex', 'eikonal2d_mex', 'migrate_mex'};
for k = 1:numel(gateways)
    mex('-largeArrayDims', ['-I' inc], ['-L' libdir], '-lsssi', ...
        ['LDFLAGS=$LDFLAGS -Wl,-rpath,' libdir], ...
        '-outdir', here, fullfile(here, [gateways{k} '.c']));
end
% the frequency-domain gateway uses the interleaved complex API
mex('-R2018a', ['-I' inc], ['-L' libdir], '-lsssi', ...
    ['LDFLAGS=$LDFLAGS -Wl,-rpath,' libdir], ...
    '-outdir', here, fullfile(here, 'freqCpmlFor2dAw_mex.c'));
